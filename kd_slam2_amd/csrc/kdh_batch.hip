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

// Pair states, forests and the per-kind tile lists of a batch, uploaded.
// mode 0: align (solve + update each iteration); mode 1: build the form once.
// dual: the CT pairs build the dual form, with the fixed velocities vf
// (n_pairs x 6).  Called with ctx->mu held and the device set.
// halves = 2: the pairs are split into two ranges of about equal tile count
// (group ct + 2 * half), for the two-stream replay of kdh_batch_align.
static int batch_setup(kdh_ctx* ctx, int n_pairs, const kdh_pair* pairs, int mode, int stats_mode, Group* g,
                       std::vector<int>& leader, int& max_iter, int dual = 0, const double* vf = nullptr,
                       int halves = 1) {
  int rc = kdh_ensure_pairs(ctx, (size_t)n_pairs);
  if (rc) return rc;
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
    PairState& ps = ctx->h_pairs[i];
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
      PairState& L = ctx->h_pairs[leader[i]];
      L.group_n += 1;
      L.group_tiles += ps.n_tiles;
    }
    max_iter = std::max(max_iter, ps.sp.max_iterations);
  }
  for (int i = 0; i < n_pairs; ++i) {
    const PairState& L = ctx->h_pairs[leader[i]];
    if (L.group_tiles == 0) {
      // no moving leaves: the first round fails the min_good gate
      PairState& ps = ctx->h_pairs[i];
      ps.iter = 1;
      ps.status = L.sp.min_good_measurements > 0 ? 2 : 1;
    }
  }
  // The split: the first forest boundary where the tiles before it reach
  // half of all tiles (no split with one half, or when nothing would follow).
  int split = n_pairs;
  if (halves == 2) {
    long long total = 0, acc = 0;
    for (int i = 0; i < n_pairs; ++i) total += ctx->h_pairs[i].n_tiles;
    for (int i = 0; i < n_pairs; ++i) {
      if (leader[i] == i && i > 0 && 2 * acc >= total) {
        split = i;
        break;
      }
      acc += ctx->h_pairs[i].n_tiles;
    }
  }
  for (int k = 0; k < kGroups; ++k) {
    g[k].p_lo = k < 2 ? 0 : split;
    g[k].p_hi = k < 2 ? split : n_pairs;
  }
  for (int i = 0; i < n_pairs; ++i) {
    PairState& ps = ctx->h_pairs[i];
    Group& G = g[ps.ct + (i >= split ? 2 : 0)];
    ps.tile_first = G.n_tiles;
    G.n_tiles += ps.n_tiles;
  }
  size_t n_partials = 0, n_tiles_all = 0;
  for (int k = 0; k < kGroups; ++k) {
    g[k].partial_off = n_partials;
    g[k].tile_off = n_tiles_all;
    n_partials += (size_t)g[k].n_tiles * slots(g[k].ct, g[k].ct ? dual : 0);
    n_tiles_all += (size_t)g[k].n_tiles;
  }
  rc = kdh_ensure_partials(ctx, std::max<size_t>(n_partials, 1));
  if (rc) return rc;
  rc = kdh_ensure_tile_pair(ctx, std::max<size_t>(n_tiles_all, 1));
  if (rc) return rc;
  std::vector<int> tp(std::max<size_t>(n_tiles_all, 1), 0);
  for (int i = 0; i < n_pairs; ++i) {
    const PairState& ps = ctx->h_pairs[i];
    const Group& G = g[ps.ct + (i >= split ? 2 : 0)];
    for (int t = 0; t < ps.n_tiles; ++t) tp[G.tile_off + ps.tile_first + t] = i;
  }
  rc = kdh_upload_tile_pair(ctx, tp);
  if (rc) return rc;
  // Only the inputs and the evolving state go up (offsetof(stats_log) of each
  // ~6 KB record): the statistics log, H, b and the dual blocks are outputs
  // the device writes before anything reads them.
  KDH_HIP_TRY(ctx, hipMemcpy2DAsync(ctx->d_pairs, sizeof(PairState), ctx->h_pairs, sizeof(PairState),
                                    kStateBytes, (size_t)n_pairs, hipMemcpyHostToDevice, ctx->stream));
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
  if (two && !ctx->stream2) {
    KDH_HIP_TRY(ctx, hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking));
    KDH_HIP_TRY(ctx, hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming));
    KDH_HIP_TRY(ctx, hipEventCreateWithFlags(&ctx->ev_stag, hipEventDisableTiming));
    KDH_HIP_TRY(ctx, hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming));
  }

  const int iters = std::min(max_iter, kMaxIter);
  // the groups of half h (kinds 0 and 1) on stream s
  auto launch_half = [&](int h, int phases, hipStream_t st) -> int {
    for (int k = 2 * h; k < 2 * h + 2; ++k) {
      if (g[k].n_tiles == 0) continue;
      int lrc = kdh_launch(ctx, ctx->d_pairs, ctx->d_tile_pair + g[k].tile_off, g[k].n_tiles, g[k].ct, nullptr,
                           nullptr, ctx->d_partials + g[k].partial_off, g[k].p_hi - g[k].p_lo, phases, 0, g[k].p_lo,
                           st);
      if (lrc != KDH_OK) return lrc;
    }
    return KDH_OK;
  };
  auto launch_phase = [&](int phases) -> int {
    int lrc = launch_half(0, phases, ctx->stream);
    return lrc != KDH_OK ? lrc : launch_half(1, phases, ctx->stream);
  };
  hipError_t le = hipSuccess;
  if (ctx->profile) {
    // Profiling: plain launches with events around each kernel phase.
    le = hipEventRecord(ctx->ev[0], ctx->stream);
    for (int it = 0; it < iters && le == hipSuccess; ++it) {
      int lrc = KDH_OK;
      for (int r = 0; r < kdh_ctx::kProfileReps && lrc == KDH_OK; ++r) lrc = launch_phase(1);
      if (lrc != KDH_OK) return lrc;
      le = hipEventRecord(ctx->ev[2 * it + 1], ctx->stream);
      if (le != hipSuccess) break;
      lrc = launch_phase(2);
      if (lrc != KDH_OK) return lrc;
      le = hipEventRecord(ctx->ev[2 * it + 2], ctx->stream);
    }
  } else {
    // Capture every iteration (tile + finish launch per non-empty group) in
    // one graph and launch it once: the pair states live in device memory,
    // so the same launches replay each Gauss-Newton step, and pairs that have
    // finished skip their tiles.  With two streams each half of the pairs is
    // its own chain of launches; the second half starts after the first
    // half's first tile launch, so from then on one half's finish stage runs
    // beside the other half's tile launch.
    const long long key[16] = {n_pairs, g[0].n_tiles, g[1].n_tiles, iters, ctx->lds_nodes, ctx->grid_override,
                               ctx->xcd_remap, (long long)(intptr_t)ctx->d_pairs,
                               (long long)(intptr_t)ctx->d_partials, (long long)(intptr_t)ctx->d_tile_pair, 2,
                               ctx->leaves_per_lane, two ? 2 : 1, g[2].n_tiles, g[3].n_tiles,
                               (long long)(intptr_t)ctx->stream2};
    hipGraphExec_t exec = kdh_graph_lookup(ctx->graph_batch, key);
    if (!exec) {
      hipGraph_t graph = nullptr;
      KDH_HIP_TRY(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      int lrc = KDH_OK;
      hipError_t ee = hipSuccess;
      if (!two) {
        for (int it = 0; it < iters && lrc == KDH_OK; ++it) lrc = launch_phase(3);
      } else {
        ee = hipEventRecord(ctx->ev_fork, ctx->stream);
        if (ee == hipSuccess) ee = hipStreamWaitEvent(ctx->stream2, ctx->ev_fork, 0);
        if (ee == hipSuccess) lrc = launch_half(0, 1, ctx->stream);
        if (ee == hipSuccess && lrc == KDH_OK) ee = hipEventRecord(ctx->ev_stag, ctx->stream);
        if (ee == hipSuccess && lrc == KDH_OK) lrc = launch_half(0, 2, ctx->stream);
        if (ee == hipSuccess) ee = hipStreamWaitEvent(ctx->stream2, ctx->ev_stag, 0);
        for (int it = 1; it < iters && lrc == KDH_OK && ee == hipSuccess; ++it) lrc = launch_half(0, 3, ctx->stream);
        for (int it = 0; it < iters && lrc == KDH_OK && ee == hipSuccess; ++it)
          lrc = launch_half(1, 3, ctx->stream2);
        if (ee == hipSuccess) ee = hipEventRecord(ctx->ev_join, ctx->stream2);
        if (ee == hipSuccess) ee = hipStreamWaitEvent(ctx->stream, ctx->ev_join, 0);
      }
      hipError_t ce = hipStreamEndCapture(ctx->stream, &graph);
      if (lrc != KDH_OK || ee != hipSuccess) {
        if (graph) hipGraphDestroy(graph);
        return lrc != KDH_OK ? lrc : kdh_hip_fail(ctx, ee, "two-stream capture");
      }
      if (ce != hipSuccess) return kdh_hip_fail(ctx, ce, "hipStreamEndCapture");
      hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      hipGraphDestroy(graph);
      if (ie != hipSuccess) return kdh_hip_fail(ctx, ie, "hipGraphInstantiate");
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
  for (int i = 0; i < n_pairs; ++i) {
    const PairState& L = ctx->h_pairs[leader[i]];
    memset(&entry[i], 0, sizeof(kdh_stats));
    if (L.iter > 0) memcpy(&entry[i], &L.stats_log[L.iter - 1], sizeof(kdh_stats));
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
    if (ps.iter > 0)
      memcpy(&r.last, &ps.stats_log[ps.iter - 1], sizeof(kdh_stats));
    else
      memset(&r.last, 0, sizeof(kdh_stats));
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
    rc = kdh_launch(ctx, ctx->d_pairs, ctx->d_tile_pair + g[k].tile_off, g[k].n_tiles, g[k].ct, nullptr, nullptr,
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
    rc = kdh_launch(ctx, ctx->d_pairs, ctx->d_tile_pair + g[1].tile_off, g[1].n_tiles, 1, nullptr, nullptr,
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
