// C ABI of the scan-alignment library (declarations and reference citations in
// include/kdh.h).  Host-side orchestration: contexts, device-resident trees,
// the per-call quadratic-form build, the device-resident Gauss-Newton loop and
// the batched launch.  No call exits the process; every HIP error becomes a
// negative return code and a message in kdh_last_error().
#include <hip/hip_runtime.h>

#include <stddef.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/kdh.h"
#include "kdh_capi_internal.h"
#include "kdh_kernels.h"
#include "kdh_solve.h"

using namespace kdh;

// ---------------------------------------------------------------------------

void kdh_params_default(kdh_params* p) {
  if (!p) return;
  p->kernel_threshold = 0.1f;
  p->damping = 1.0f;
  p->max_error = 1.0f;
  p->min_normal_cos = (float)cos(25. / 180. * M_PI);
  p->min_good_measurements = 10;
  p->max_iterations = 20;
  p->velocity_damping = 100.0f;
  p->use_term_crit = 0;
  p->term_d_pert = 1e-3f;
  p->term_d_score = 1e-3f;
  p->recenter = 1;
}

int kdh_fail(kdh_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return code;
}

int kdh_hip_fail(kdh_ctx* ctx, hipError_t e, const char* where) {
  std::string m = std::string(where) + ": " + hipGetErrorString(e);
  return kdh_fail(ctx, KDH_ERR_HIP, m);
}

#define KDH_HIP_TRY(ctx, expr)                              \
  do {                                                      \
    hipError_t _e = (expr);                                 \
    if (_e != hipSuccess) return kdh_hip_fail(ctx, _e, #expr); \
  } while (0)

int kdh_ctx_create(int device, kdh_ctx** out) {
  if (!out) return KDH_ERR_INVALID;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return KDH_ERR_HIP;
  if (device < 0 || device >= n) return KDH_ERR_INVALID;
  if (hipSetDevice(device) != hipSuccess) return KDH_ERR_HIP;
  kdh_ctx* c = new kdh_ctx();
  c->device = device;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return KDH_ERR_HIP;
  }
  // coherent: kernels write results into it (k_pull_results) for the host
  if (hipHostMalloc((void**)&c->h_pairs, sizeof(PairState) * 1, hipHostMallocCoherent) != hipSuccess) {
    hipStreamDestroy(c->stream);
    delete c;
    return KDH_ERR_NOMEM;
  }
  if (hipHostGetDevicePointer((void**)&c->h_pairs_dev, c->h_pairs, 0) != hipSuccess || !c->h_pairs_dev) {
    hipHostFree(c->h_pairs);
    hipStreamDestroy(c->stream);
    delete c;
    return KDH_ERR_HIP;
  }
  c->h_pairs_cap = 1;
  for (int i = 0; i < kdh_ctx::kEvents; ++i) {
    if (hipEventCreate(&c->ev[i]) != hipSuccess) {
      delete c;
      return KDH_ERR_HIP;
    }
  }
  *out = c;
  return KDH_OK;
}

int kdh_ctx_destroy(kdh_ctx* ctx) {
  if (!ctx) return KDH_ERR_INVALID;
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  if (ctx->graph_align.exec) hipGraphExecDestroy(ctx->graph_align.exec);
  if (ctx->graph_batch.exec) hipGraphExecDestroy(ctx->graph_batch.exec);
  if (ctx->d_pairs) hipFree(ctx->d_pairs);
  if (ctx->d_partials) hipFree(ctx->d_partials);
  if (ctx->d_tiles) hipFree(ctx->d_tiles);
  if (ctx->d_corr) hipFree(ctx->d_corr);
  if (ctx->d_depth) hipFree(ctx->d_depth);
  if (ctx->d_extract) hipFree(ctx->d_extract);
  if (ctx->h_root) hipHostFree(ctx->h_root);
  kdh_tree_build_ws_free(ctx);
  if (ctx->h_pairs) hipHostFree(ctx->h_pairs);
  if (ctx->h_tiles) hipHostFree(ctx->h_tiles);
  for (int i = 0; i < kdh_ctx::kEvents; ++i)
    if (ctx->ev[i]) hipEventDestroy(ctx->ev[i]);
  for (hipEvent_t e : {ctx->ev_fork, ctx->ev_stag, ctx->ev_join})
    if (e) hipEventDestroy(e);
  if (ctx->stream2) {
    hipStreamSynchronize(ctx->stream2);
    hipStreamDestroy(ctx->stream2);
  }
  hipStreamDestroy(ctx->stream);
  delete ctx;
  return KDH_OK;
}

const char* kdh_last_error(const kdh_ctx* ctx) {
  if (!ctx) return "null context";
  return ctx->err.c_str();
}

int kdh_set_launch_config(kdh_ctx* ctx, int grid, int lds_nodes) {
  if (!ctx) return KDH_ERR_INVALID;
  if (grid < 0) return kdh_fail(ctx, KDH_ERR_INVALID, "grid must be >= 0");
  if (lds_nodes != 0 && lds_nodes != 256 && lds_nodes != 512 && lds_nodes != 1024 && lds_nodes != 2048)
    return kdh_fail(ctx, KDH_ERR_INVALID, "lds_nodes must be one of 0, 256, 512, 1024, 2048");
  std::lock_guard<std::mutex> lk(ctx->mu);
  ctx->grid_override = grid;
  ctx->lds_nodes = lds_nodes;
  return KDH_OK;
}

int kdh_set_batch_streams(kdh_ctx* ctx, int streams) {
  if (!ctx) return KDH_ERR_INVALID;
  if (streams != 1 && streams != 2) return kdh_fail(ctx, KDH_ERR_INVALID, "batch streams must be 1 or 2");
  std::lock_guard<std::mutex> lk(ctx->mu);
  ctx->batch_streams = streams;
  return KDH_OK;
}

int kdh_set_leaves_per_lane(kdh_ctx* ctx, int leaves_per_lane) {
  if (!ctx) return KDH_ERR_INVALID;
  if (leaves_per_lane != 1 && leaves_per_lane != 2)
    return kdh_fail(ctx, KDH_ERR_INVALID, "leaves_per_lane must be 1 or 2");
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (leaves_per_lane == 2 && ctx->tile_leaves != 256)
    return kdh_fail(ctx, KDH_ERR_INVALID, "two leaves per lane take 256-leaf tiles");
  ctx->leaves_per_lane = leaves_per_lane;
  return KDH_OK;
}

int kdh_set_tile_leaves(kdh_ctx* ctx, int tile_leaves) {
  if (!ctx) return KDH_ERR_INVALID;
  if (tile_leaves != 256 && tile_leaves != 128)
    return kdh_fail(ctx, KDH_ERR_INVALID, "tile_leaves must be 256 or 128");
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (tile_leaves == 128 && ctx->leaves_per_lane != 1)
    return kdh_fail(ctx, KDH_ERR_INVALID, "128-leaf tiles take one leaf per lane");
  ctx->tile_leaves = tile_leaves;
  return KDH_OK;
}

// ---- trees -----------------------------------------------------------------

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

int kdh_tree_upload(kdh_ctx* ctx, const kdh_node3f* nodes, int64_t n_nodes, const uint64_t* leaf_indices,
                    int64_t n_leaves, kdh_tree** out) {
  if (!ctx || !out) return KDH_ERR_INVALID;
  *out = nullptr;
  if (n_nodes < 0 || n_leaves < 0 || (n_nodes > 0 && !nodes) || (n_leaves > 0 && !leaf_indices))
    return kdh_fail(ctx, KDH_ERR_INVALID, "kdh_tree_upload: bad sizes or null arrays");
  if (n_nodes > (int64_t)0x3fffffff || n_leaves > n_nodes)
    return kdh_fail(ctx, KDH_ERR_INVALID, "kdh_tree_upload: tree too large or more leaves than nodes");
  // Validate links and leaf indices on the host: the kernels index without checks.
  for (int64_t i = 0; i < n_nodes; ++i) {
    const int32_t l = nodes[i].idx_left, r = nodes[i].idx_right;
    if (l >= (int32_t)n_nodes || r >= (int32_t)n_nodes || l < -1 || r < -1)
      return kdh_fail(ctx, KDH_ERR_INVALID, "kdh_tree_upload: child link out of range at node " + std::to_string(i));
    if ((l >= 0 && l <= i) || (r >= 0 && r <= i))
      return kdh_fail(ctx, KDH_ERR_INVALID,
                      "kdh_tree_upload: child link does not point forward (tree not in BFS order) at node " +
                          std::to_string(i));
  }
  for (int64_t i = 0; i < n_leaves; ++i) {
    if (leaf_indices[i] >= (uint64_t)n_nodes)
      return kdh_fail(ctx, KDH_ERR_INVALID, "kdh_tree_upload: leaf index out of range");
  }
  const size_t node_bytes = align256((size_t)std::max<int64_t>(n_nodes, 1) * 32);
  const size_t leaf_bytes = align256((size_t)std::max<int64_t>(n_leaves, 1) * 32);
  const size_t dts_bytes = align256((size_t)std::max<int64_t>(n_nodes, 1) * 4);
  std::vector<float> staging((node_bytes + leaf_bytes + dts_bytes) / 4, 0.0f);
  float* sn = staging.data();
  for (int64_t i = 0; i < n_nodes; ++i) {
    const kdh_node3f& n = nodes[i];
    float* o = sn + i * 8;
    o[0] = n.mean[0]; o[1] = n.mean[1]; o[2] = n.mean[2]; o[3] = n.direction[0];
    o[4] = n.direction[1]; o[5] = n.direction[2];
    memcpy(&o[6], &n.idx_left, 4);
    memcpy(&o[7], &n.idx_right, 4);
  }
  float* sl = staging.data() + node_bytes / 4;
  for (int64_t i = 0; i < n_leaves; ++i) {
    const kdh_node3f& n = nodes[leaf_indices[i]];
    float* o = sl + i * 8;
    o[0] = n.mean[0]; o[1] = n.mean[1]; o[2] = n.mean[2]; o[3] = n.direction[0];
    o[4] = n.direction[1]; o[5] = n.direction[2];
    o[6] = (float)n.dts_mean;  // narrowed as `Scalar dt = _dts_mean` (cticp_impl_.h:158)
    int32_t idx = (int32_t)leaf_indices[i];
    memcpy(&o[7], &idx, 4);
  }
  float* sd = staging.data() + (node_bytes + leaf_bytes) / 4;
  for (int64_t i = 0; i < n_nodes; ++i) sd[i] = (float)nodes[i].dts_mean;  // Scalar_(dts_mean)
  std::lock_guard<std::mutex> lk(ctx->mu);
  KDH_HIP_TRY(ctx, hipSetDevice(ctx->device));
  void* d = nullptr;
  const size_t total = node_bytes + leaf_bytes + dts_bytes;
  hipError_t e = hipMalloc(&d, total);
  if (e != hipSuccess) return kdh_fail(ctx, KDH_ERR_NOMEM, std::string("kdh_tree_upload: hipMalloc: ") + hipGetErrorString(e));
  e = hipMemcpyAsync(d, staging.data(), total, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    hipFree(d);
    return kdh_hip_fail(ctx, e, "kdh_tree_upload: copy");
  }
  kdh_tree* t = new kdh_tree();
  t->device = ctx->device;
  t->dmem = d;
  t->bytes = total;
  t->dt.nodes = (const float4*)d;
  t->dt.leaves = (const float4*)((char*)d + node_bytes);
  t->node_dts = (float*)((char*)d + node_bytes + leaf_bytes);
  t->dt.node_dts = t->node_dts;
  t->dt.n_nodes = (int)n_nodes;
  t->dt.n_leaves = (int)n_leaves;
  t->root_mean = n_nodes > 0 ? v3(nodes[0].mean[0], nodes[0].mean[1], nodes[0].mean[2]) : v3(0.0f, 0.0f, 0.0f);
  *out = t;
  return KDH_OK;
}

int kdh_tree_free(kdh_ctx* ctx, kdh_tree* tree) {
  if (!ctx || !tree) return KDH_ERR_INVALID;
  std::lock_guard<std::mutex> lk(ctx->mu);
  hipSetDevice(tree->device);
  hipStreamSynchronize(ctx->stream);
  hipError_t e = hipFree(tree->dmem);
  delete tree;
  if (e != hipSuccess) return kdh_hip_fail(ctx, e, "kdh_tree_free");
  return KDH_OK;
}

int kdh_tree_info(const kdh_tree* tree, int64_t* n_nodes, int64_t* n_leaves, int64_t* device_bytes) {
  if (!tree) return KDH_ERR_INVALID;
  if (n_nodes) *n_nodes = tree->dt.n_nodes;
  if (n_leaves) *n_leaves = tree->dt.n_leaves;
  if (device_bytes) *device_bytes = (int64_t)tree->bytes;
  return KDH_OK;
}

// ---- workspace ---------------------------------------------------------------

template <typename T>
static hipError_t grow(T** p, size_t* cap, size_t need) {
  if (need <= *cap && *p) return hipSuccess;
  size_t nc = std::max<size_t>(need, 2 * *cap);
  if (*p) {
    hipError_t e = hipFree(*p);
    if (e != hipSuccess) return e;
    *p = nullptr;
  }
  hipError_t e = hipMalloc((void**)p, nc * sizeof(T));
  if (e == hipSuccess) *cap = nc;
  else *cap = 0;
  return e;
}

int kdh_ensure_pairs(kdh_ctx* ctx, size_t n_pairs) {
  KDH_HIP_TRY(ctx, grow(&ctx->d_pairs, &ctx->pairs_cap, n_pairs));
  if (n_pairs > ctx->h_pairs_cap) {
    if (ctx->h_pairs) hipHostFree(ctx->h_pairs);
    ctx->h_pairs = nullptr;
    size_t nc = std::max(n_pairs, 2 * ctx->h_pairs_cap);
    // coherent, as at context creation: kernels write results into it
    hipError_t e = hipHostMalloc((void**)&ctx->h_pairs, nc * sizeof(PairState), hipHostMallocCoherent);
    if (e == hipSuccess) e = hipHostGetDevicePointer((void**)&ctx->h_pairs_dev, ctx->h_pairs, 0);
    if (e != hipSuccess) {
      if (ctx->h_pairs) hipHostFree(ctx->h_pairs);
      ctx->h_pairs = ctx->h_pairs_dev = nullptr;
      ctx->h_pairs_cap = 0;
      return kdh_fail(ctx, KDH_ERR_NOMEM, "pinned pair buffer");
    }
    ctx->h_pairs_cap = nc;
  }
  return KDH_OK;
}

// One workgroup per pair writes its results straight into the pinned
// mirror.  A kernel rather than a DMA copy: the DMA path of a 200 KB
// device-to-host copy stalled for about 5 ms about once in 200 served steps.
__global__ void k_pull_results(const PairState* __restrict__ d, PairState* h, int n) {
  const int p = blockIdx.x, t = threadIdx.x;
  if (p >= n) return;
  const PairState& s = d[p];
  PairState& o = h[p];
  if (t < 12) o.X[t] = s.X[t];
  if (t < 6) o.v[t] = s.v[t];
  if (t == 0) {
    o.iter = s.iter;
    o.status = s.status;
    o.prev_score = s.prev_score;
  }
  static_assert(sizeof(DevStats) % 8 == 0, "stats copied as doubles");
  const int words = (int)(sizeof(DevStats) / 8) * min(max(s.iter, 0), kMaxIter);
  const double* src = reinterpret_cast<const double*>(s.stats_log);
  double* dst = reinterpret_cast<double*>(o.stats_log);
  for (int i = t; i < words; i += 64) dst[i] = src[i];
  // no fence of its own: the host reads the mirror after synchronising with
  // the stream, and the kernel's completion makes these writes visible to it
}

int kdh_pull_results(kdh_ctx* ctx, int n_pairs) {
  if (n_pairs <= 0) return KDH_OK;
  if (!ctx->h_pairs_dev || (size_t)n_pairs > ctx->h_pairs_cap || (size_t)n_pairs > ctx->pairs_cap)
    return kdh_fail(ctx, KDH_ERR_INVALID, "kdh_pull_results: pinned mirror not mapped or too small");
  hipLaunchKernelGGL(k_pull_results, dim3(n_pairs), dim3(64), 0, ctx->stream, ctx->d_pairs, ctx->h_pairs_dev, n_pairs);
  KDH_HIP_TRY(ctx, hipGetLastError());
  return KDH_OK;
}

// Build mode's outputs (H, b, the dual blocks when asked, the statistics
// record) written into the pinned mirror the same way.
__global__ void k_pull_forms(const PairState* __restrict__ d, PairState* h, int n, int dual) {
  const int p = blockIdx.x, t = threadIdx.x;
  if (p >= n) return;
  const PairState& s = d[p];
  PairState& o = h[p];
  for (int i = t; i < 144; i += 64) o.H[i] = s.H[i];
  if (t < 12) o.b[t] = s.b[t];
  if (dual)
    for (int i = t; i < 114; i += 64) o.Hd[i] = s.Hd[i];
  const double* src = reinterpret_cast<const double*>(&s.stats_log[0]);
  double* dst = reinterpret_cast<double*>(&o.stats_log[0]);
  for (int i = t; i < (int)(sizeof(DevStats) / 8); i += 64) dst[i] = src[i];
  // visible to the host after its stream synchronisation, as k_pull_results
}

int kdh_pull_forms(kdh_ctx* ctx, int n_pairs, int dual) {
  if (n_pairs <= 0) return KDH_OK;
  if (!ctx->h_pairs_dev || (size_t)n_pairs > ctx->h_pairs_cap || (size_t)n_pairs > ctx->pairs_cap)
    return kdh_fail(ctx, KDH_ERR_INVALID, "kdh_pull_forms: pinned mirror not mapped or too small");
  hipLaunchKernelGGL(k_pull_forms, dim3(n_pairs), dim3(64), 0, ctx->stream, ctx->d_pairs, ctx->h_pairs_dev, n_pairs,
                     dual);
  KDH_HIP_TRY(ctx, hipGetLastError());
  return KDH_OK;
}

int kdh_ensure_partials(kdh_ctx* ctx, size_t n_floats) {
  KDH_HIP_TRY(ctx, grow(&ctx->d_partials, &ctx->partials_cap, n_floats));
  return KDH_OK;
}

int kdh_ensure_tiles(kdh_ctx* ctx, size_t n) {
  TileDesc* before = ctx->d_tiles;
  KDH_HIP_TRY(ctx, grow(&ctx->d_tiles, &ctx->tile_cap, n));
  if (ctx->d_tiles != before) ctx->tiles_on_device.clear();
  return KDH_OK;
}

bool kdh_same_tiles(const std::vector<TileDesc>& a, const std::vector<TileDesc>& b) {
  return a.size() == b.size() && (a.empty() || memcmp(a.data(), b.data(), a.size() * sizeof(TileDesc)) == 0);
}

int kdh_upload_tiles(kdh_ctx* ctx, const std::vector<TileDesc>& tiles) {
  if (kdh_same_tiles(tiles, ctx->tiles_on_device)) return KDH_OK;
  if (tiles.size() > ctx->h_tile_cap) {
    if (ctx->h_tiles) hipHostFree(ctx->h_tiles);
    ctx->h_tiles = nullptr;
    ctx->h_tile_cap = 0;
    const size_t nc = std::max(tiles.size(), (size_t)4096);
    KDH_HIP_TRY(ctx, hipHostMalloc((void**)&ctx->h_tiles, nc * sizeof(TileDesc), hipHostMallocDefault));
    ctx->h_tile_cap = nc;
  }
  // the staging buffer may still feed an earlier copy on the stream
  KDH_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  memcpy(ctx->h_tiles, tiles.data(), tiles.size() * sizeof(TileDesc));
  KDH_HIP_TRY(ctx, hipMemcpyAsync(ctx->d_tiles, ctx->h_tiles, tiles.size() * sizeof(TileDesc), hipMemcpyHostToDevice,
                                  ctx->stream));
  ctx->tiles_on_device = tiles;
  return KDH_OK;
}

// ---- pair set-up ---------------------------------------------------------------

static TermParams term_params(const kdh_params& p, int stats_mode) {
  TermParams t;
  t.kernel_threshold = p.kernel_threshold;
  t.max_error = p.max_error;
  t.min_normal_cos = p.min_normal_cos;
  t.stats_mode = stats_mode ? 1 : 0;
  t.recenter = p.recenter ? 1 : 0;
  t.kt_quarter = p.kernel_threshold > 0.0f ? (float)sqrt(sqrt((double)p.kernel_threshold)) : 0.0f;
  return t;
}

static SolveParams solve_params(const kdh_params& p) {
  SolveParams s;
  s.damping = p.damping;
  s.velocity_damping = p.velocity_damping;
  s.min_good_measurements = p.min_good_measurements;
  s.max_iterations = p.max_iterations;
  s.use_term_crit = p.use_term_crit;
  s.d_pert = p.term_d_pert;
  s.d_score = p.term_d_score;
  s.recenter = p.recenter;
  return s;
}

static int n_tiles_of(int n_leaves, int tile_leaves) { return (n_leaves + tile_leaves - 1) / tile_leaves; }

// Initial FP32 cache from FP64 state: the same formula the device applies
// after every update (kdh_kernels.hip, finish_pair).
void kdh_cache_from_state(PairState& ps) { cache_from_state(ps); }

int kdh_fill_pair(kdh_ctx* ctx, PairState& ps, const kdh_tree* fixed, const kdh_tree* moving, int ct,
                  const kdh_params& p, int stats_mode, int mode) {
  if (!fixed || !moving) return kdh_fail(ctx, KDH_ERR_INVALID, "null tree");
  if (fixed->device != ctx->device || moving->device != ctx->device)
    return kdh_fail(ctx, KDH_ERR_INVALID, "tree lives on another device");
  if (mode == 0 && (p.max_iterations < 1 || p.max_iterations > kMaxIter))
    return kdh_fail(ctx, KDH_ERR_INVALID, "max_iterations must be in [1, 64] for the device loop");
  memset(&ps, 0, offsetof(PairState, stats_log));
  ps.fixed = fixed->dt;
  ps.moving = moving->dt;
  ps.tp = term_params(p, stats_mode);
  ps.sp = solve_params(p);
  if (stats_mode) ps.sp.max_iterations = 1;
  ps.ct = ct ? 1 : 0;
  ps.mode = mode;
  ps.n_tiles = n_tiles_of(moving->dt.n_leaves, ctx->tile_leaves);
  ps.group_n = 1;
  ps.group_tiles = ps.n_tiles;
  ps.moving_root_mean = moving->root_mean;
  ps.status = 0;
  ps.iter = 0;
  return KDH_OK;
}

int kdh_launch(kdh_ctx* ctx, PairState* d_pairs, const TileDesc* d_tiles, int n_tiles, int ct, int* corr,
               int* depth, float* partials, int n_pairs, int phases, int dual, int pair_first, hipStream_t stream) {
  LaunchArgs a;
  a.dual = dual;
  a.pair_first = pair_first;
  a.n_pairs = n_pairs;
  a.phases = phases;
  a.pair_tile_first = nullptr;
  a.status_out = nullptr;
  a.pairs = d_pairs;
  a.tiles = d_tiles;
  a.n_tiles_total = n_tiles;
  a.partials = partials ? partials : ctx->d_partials;
  a.xcd_remap = ctx->xcd_remap;
  a.corr_out = corr;
  a.depth_out = depth;
  a.grid = kdh_tile_grid(ctx, n_tiles);
  a.ct = ct;
  int lds = ctx->lds_nodes;
  if (ct && lds > 1024) lds = 1024;
  if (dual && lds > 512) lds = 512;
  a.lds_nodes = lds;
  a.leaves_per_lane = dual ? 1 : ctx->leaves_per_lane;
  a.tile_leaves = ctx->tile_leaves;
  if (n_tiles <= 0) return KDH_OK;
  hipError_t e = launch_tiles(a, stream ? stream : ctx->stream);
  if (e != hipSuccess) return kdh_hip_fail(ctx, e, "tile kernel launch");
  return KDH_OK;
}

// ---- per-call build ---------------------------------------------------------------

// The inputs of one build: the reference's FP32 cache (Xm, Xf, Xmf, v), or an
// FP64 state (X, Xf, v) the cache is rounded from.
struct BuildInput {
  const float *Xm = nullptr, *Xf = nullptr, *Xmf = nullptr, *v = nullptr;
  const double *X64 = nullptr, *Xf64 = nullptr, *v64 = nullptr;
};

static int build_common(kdh_ctx* ctx, const kdh_tree* fixed, const kdh_tree* moving, int ct, const BuildInput& in,
                        const kdh_params* params, int stats_mode, double* H, double* b, kdh_stats* stats,
                        int32_t* leaf_out, int32_t* depth_out) {
  if (!ctx) return KDH_ERR_INVALID;
  const bool at_state = in.X64 != nullptr;
  if (!params || (at_state ? (!in.Xf64 || (ct && !in.v64)) : (!in.Xm || !in.Xf || !in.Xmf || (ct && !in.v))))
    return kdh_fail(ctx, KDH_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(ctx->mu);
  KDH_HIP_TRY(ctx, hipSetDevice(ctx->device));
  int rc = kdh_ensure_pairs(ctx, 1);
  if (rc) return rc;
  PairState& ps = ctx->h_pairs[0];
  rc = kdh_fill_pair(ctx, ps, fixed, moving, ct, *params, stats_mode, 1);
  if (rc) return rc;
  if (at_state) {
    for (int i = 0; i < 12; ++i) {
      ps.X[i] = in.X64[i];
      ps.Xf[i] = in.Xf64[i];
    }
    for (int i = 0; i < 6; ++i) ps.v[i] = ct ? in.v64[i] : 0.0;
    kdh_cache_from_state(ps);
  } else {
    ps.cache.Xm = iso_from_34<float>(in.Xm);
    ps.cache.Xf = iso_from_34<float>(in.Xf);
    ps.cache.Xmf = iso_from_34<float>(in.Xmf);
    for (int i = 0; i < 6; ++i) ps.cache.v[i] = ct ? in.v[i] : 0.0f;
    // The caller's FP32 cache is the state here.
    set_anchor(ps.cache, ps.c64,
               make_anchor(iso_cast<double>(ps.cache.Xm), iso_cast<double>(ps.cache.Xf), moving->root_mean,
                           params->recenter));
  }
  const int nt = ps.n_tiles;
  rc = kdh_ensure_partials(ctx, (size_t)std::max(nt, 1) * slots(ct));
  if (rc) return rc;
  const size_t nl = (size_t)moving->dt.n_leaves;
  if (leaf_out) {
    KDH_HIP_TRY(ctx, grow(&ctx->d_corr, &ctx->corr_cap, std::max<size_t>(nl, 1)));
    KDH_HIP_TRY(ctx, grow(&ctx->d_depth, &ctx->depth_cap, std::max<size_t>(nl, 1)));
  }
  KDH_HIP_TRY(ctx, hipMemcpyAsync(ctx->d_pairs, &ps, offsetof(PairState, stats_log), hipMemcpyHostToDevice,
                                  ctx->stream));
  rc = kdh_launch(ctx, ctx->d_pairs, nullptr, nt, ct, leaf_out ? ctx->d_corr : nullptr,
                  leaf_out ? ctx->d_depth : nullptr, nullptr, 1);
  if (rc) return rc;
  if (nt > 0) {
    KDH_HIP_TRY(ctx, hipMemcpyAsync(&ps.stats_log[0], &ctx->d_pairs->stats_log[0], sizeof(DevStats),
                                    hipMemcpyDeviceToHost, ctx->stream));
    KDH_HIP_TRY(ctx, hipMemcpyAsync(ps.H, ctx->d_pairs->H, sizeof(ps.H) + sizeof(ps.b), hipMemcpyDeviceToHost,
                                    ctx->stream));
  }
  std::vector<int32_t> hc, hd;
  if (leaf_out && nl > 0) {
    KDH_HIP_TRY(ctx, hipMemcpyAsync(leaf_out, ctx->d_corr, nl * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (depth_out)
      KDH_HIP_TRY(ctx, hipMemcpyAsync(depth_out, ctx->d_depth, nl * 4, hipMemcpyDeviceToHost, ctx->stream));
  }
  KDH_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  const int D = ct ? 12 : 6;
  if (nt == 0) {
    // empty moving tree: nothing to reduce
    memset(ps.H, 0, sizeof(ps.H));
    memset(ps.b, 0, sizeof(ps.b));
    memset(&ps.stats_log[0], 0, sizeof(DevStats));
  }
  if (H) memcpy(H, ps.H, sizeof(double) * D * D);
  if (b) memcpy(b, ps.b, sizeof(double) * D);
  if (stats) memcpy(stats, &ps.stats_log[0], sizeof(kdh_stats));
  return KDH_OK;
}

int kdh_icp_build(kdh_ctx* ctx, const kdh_tree* fixed, const kdh_tree* moving, const float Xm[12], const float Xf[12],
                  const float Xmf[12], const kdh_params* params, int stats_mode, double H[36], double b[6],
                  kdh_stats* stats) {
  BuildInput in;
  in.Xm = Xm;
  in.Xf = Xf;
  in.Xmf = Xmf;
  return build_common(ctx, fixed, moving, 0, in, params, stats_mode, H, b, stats, nullptr, nullptr);
}

int kdh_cticp_build(kdh_ctx* ctx, const kdh_tree* fixed, const kdh_tree* moving, const float Xm[12],
                    const float Xf[12], const float Xmf[12], const float v[6], const kdh_params* params,
                    int stats_mode, double H[144], double b[12], kdh_stats* stats) {
  BuildInput in;
  in.Xm = Xm;
  in.Xf = Xf;
  in.Xmf = Xmf;
  in.v = v;
  return build_common(ctx, fixed, moving, 1, in, params, stats_mode, H, b, stats, nullptr, nullptr);
}

int kdh_build_at(kdh_ctx* ctx, const kdh_tree* fixed, const kdh_tree* moving, int ct, const double X[12],
                 const double Xf[12], const double v[6], const kdh_params* params, int stats_mode, double* H,
                 double* b, kdh_stats* stats) {
  BuildInput in;
  in.X64 = X;
  in.Xf64 = Xf;
  in.v64 = v;
  return build_common(ctx, fixed, moving, ct ? 1 : 0, in, params, stats_mode, H, b, stats, nullptr, nullptr);
}

int kdh_correspondences(kdh_ctx* ctx, const kdh_tree* fixed, const kdh_tree* moving, int ct, const float Xm[12],
                        const float Xf[12], const float Xmf[12], const float v[6], const kdh_params* params,
                        int32_t* leaf_out, int32_t* depth_out) {
  if (!leaf_out) return kdh_fail(ctx, KDH_ERR_INVALID, "leaf_out is null");
  float zero[6] = {0, 0, 0, 0, 0, 0};
  BuildInput in;
  in.Xm = Xm;
  in.Xf = Xf;
  in.Xmf = Xmf;
  in.v = v ? v : zero;
  return build_common(ctx, fixed, moving, ct, in, params, 0, nullptr, nullptr, nullptr, leaf_out, depth_out);
}

// ---- whole alignment on the device ---------------------------------------------------

// compute()'s return value (tree_aligner_base_.h:100-115): -i when round i
// failed the min_good gate, i when the criterion stopped after round i,
// otherwise the configured max_iterations (also in stats mode).
int kdh_result_code_of(int status, int iter, int max_iterations_param) {
  if (status == 2) return -(iter - 1);
  if (status == 3) return iter - 1;
  return max_iterations_param;
}

int kdh_result_code(const PairState& ps, int max_iterations_param) {
  return kdh_result_code_of(ps.status, ps.iter, max_iterations_param);
}

int kdh_setup_prior(PairState& ps, const kdh_prior* prior) {
  ps.has_prior = 0.0;
  if (!prior) return KDH_OK;
  bool nz = false;
  for (int i = 0; i < 36; ++i) nz = nz || prior->info[i] != 0.0;
  if (!nz) return KDH_OK;
  IsoD Xr = iso_from_34<double>(prior->X_ref);
  IsoD Z = iso_from_34<double>(prior->Z);
  IsoD ZXi = inverse(compose(Xr, Z));
  iso_to_34(ZXi, ps.prior_ZXi);
  memcpy(ps.prior_info, prior->info, sizeof(ps.prior_info));
  ps.has_prior = 1.0;
  return KDH_OK;
}

int kdh_align(kdh_ctx* ctx, const kdh_tree* fixed, const kdh_tree* moving, int ct, const double X0[12],
              const double v0[6], const double Xf[12], const kdh_params* params, const kdh_prior* prior,
              int stats_mode, double X_out[12], double v_out[6], kdh_stats* stats_log, int* n_log, int* result,
              float* cache_log) {
  if (!ctx) return KDH_ERR_INVALID;
  if (!X0 || !Xf || !params) return kdh_fail(ctx, KDH_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(ctx->mu);
  KDH_HIP_TRY(ctx, hipSetDevice(ctx->device));
  int rc = kdh_ensure_pairs(ctx, 1);
  if (rc) return rc;
  PairState& ps = ctx->h_pairs[0];
  rc = kdh_fill_pair(ctx, ps, fixed, moving, ct, *params, stats_mode, 0);
  if (rc) return rc;
  memcpy(ps.X, X0, sizeof(ps.X));
  memcpy(ps.Xf, Xf, sizeof(ps.Xf));
  for (int i = 0; i < 6; ++i) ps.v[i] = (ct && v0) ? v0[i] : 0.0;
  kdh_setup_prior(ps, prior);
  kdh_cache_from_state(ps);
  const int nt = ps.n_tiles;
  const int iters = ps.sp.max_iterations;
  rc = kdh_ensure_partials(ctx, (size_t)std::max(nt, 1) * slots(ct));
  if (rc) return rc;
  KDH_HIP_TRY(ctx, hipMemcpyAsync(ctx->d_pairs, &ps, sizeof(PairState), hipMemcpyHostToDevice, ctx->stream));
  if (nt == 0) {
    // No moving leaves: every round fails the min_good gate, as the reference does.
    KDH_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    memset(&ps.stats_log[0], 0, sizeof(DevStats));
    ps.iter = 1;
    ps.status = params->min_good_measurements > 0 ? 2 : 1;
  } else {
    const bool prof = ctx->profile != 0;
    if (!prof && !cache_log) {
      // Production path: every iteration captured in one graph, one launch;
      // the executable graph is cached while the launch shape is unchanged.
      const long long key[16] = {nt, ct, iters, ctx->lds_nodes, ctx->grid_override, ctx->xcd_remap,
                                 (long long)(intptr_t)ctx->d_pairs, (long long)(intptr_t)ctx->d_partials, 1,
                                 ctx->leaves_per_lane, ctx->tile_leaves, 0, 0, 0, 0, 0};
      hipGraphExec_t exec = kdh_graph_lookup(ctx->graph_align, key);
      if (!exec) {
        hipGraph_t graph = nullptr;
        KDH_HIP_TRY(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        int lrc = KDH_OK;
        for (int it = 0; it < iters && lrc == KDH_OK; ++it)
          lrc = kdh_launch(ctx, ctx->d_pairs, nullptr, nt, ct, nullptr, nullptr, nullptr, 1, 3);
        hipError_t ce = hipStreamEndCapture(ctx->stream, &graph);
        if (lrc != KDH_OK) {
          if (graph) hipGraphDestroy(graph);
          return lrc;
        }
        if (ce != hipSuccess) return kdh_hip_fail(ctx, ce, "hipStreamEndCapture");
        hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        if (ie != hipSuccess) return kdh_hip_fail(ctx, ie, "hipGraphInstantiate");
        kdh_graph_store(ctx->graph_align, key, exec);
      }
      hipError_t le = hipEventRecord(ctx->ev[0], ctx->stream);
      if (le == hipSuccess) le = hipGraphLaunch(exec, ctx->stream);
      if (le == hipSuccess) le = hipEventRecord(ctx->ev[1], ctx->stream);
      if (le == hipSuccess) {
        rc = kdh_pull_results(ctx, 1);  // into ps, the pinned mirror's first state
        if (rc) return rc;
        le = hipStreamSynchronize(ctx->stream);
      }
      if (le != hipSuccess) return kdh_hip_fail(ctx, le, "device loop graph");
      float ms = 0.0f;
      hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]);
      ctx->last_kernel_us = 1000.0 * ms / iters;
      ctx->last_tile_us = ctx->last_finish_us = 0.0;
      ctx->last_launches = iters;
    } else {
    KDH_HIP_TRY(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
    for (int it = 0; it < iters; ++it) {
      if (cache_log) {
        KDH_HIP_TRY(ctx, hipMemcpyAsync(&ps.cache, &ctx->d_pairs->cache, sizeof(TermCache), hipMemcpyDeviceToHost,
                                        ctx->stream));
        KDH_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        iso_to_34(ps.cache.Xm, cache_log + it * 36);
        iso_to_34(ps.cache.Xf, cache_log + it * 36 + 12);
        iso_to_34(ps.cache.Xmf, cache_log + it * 36 + 24);
      }
      if (prof) {
        // The tile kernel is idempotent (it only rewrites the partial rows),
        // so it runs kProfileReps times back to back between the events.
        for (int r = 0; r < kdh_ctx::kProfileReps; ++r) {
          rc = kdh_launch(ctx, ctx->d_pairs, nullptr, nt, ct, nullptr, nullptr, nullptr, 1, 1);
          if (rc) return rc;
        }
        KDH_HIP_TRY(ctx, hipEventRecord(ctx->ev[2 * it + 1], ctx->stream));
        rc = kdh_launch(ctx, ctx->d_pairs, nullptr, nt, ct, nullptr, nullptr, nullptr, 1, 2);
        if (rc) return rc;
        KDH_HIP_TRY(ctx, hipEventRecord(ctx->ev[2 * it + 2], ctx->stream));
      } else {
        rc = kdh_launch(ctx, ctx->d_pairs, nullptr, nt, ct, nullptr, nullptr, nullptr, 1, 3);
        if (rc) return rc;
        KDH_HIP_TRY(ctx, hipEventRecord(ctx->ev[2 * it + 2], ctx->stream));
      }
    }
    KDH_HIP_TRY(ctx, hipMemcpyAsync(&ps, ctx->d_pairs, sizeof(PairState), hipMemcpyDeviceToHost, ctx->stream));
    KDH_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    kdh_collect_times(ctx, iters, prof);
    }
  }
  if (X_out) memcpy(X_out, ps.X, sizeof(ps.X));
  if (v_out) memcpy(v_out, ps.v, sizeof(ps.v));
  const int n = ps.iter;
  if (stats_log) memcpy(stats_log, ps.stats_log, sizeof(DevStats) * n);
  if (n_log) *n_log = n;
  if (result) *result = kdh_result_code(ps, params->max_iterations);
  return KDH_OK;
}

// Event layout per iteration it: ev[2it] .. ev[2it+2] the whole iteration;
// in profiling mode ev[2it] .. ev[2it+1] brackets kProfileReps back-to-back
// tile kernels and ev[2it+1] .. ev[2it+2] the finish kernel.
void kdh_collect_times(kdh_ctx* ctx, int iters, bool prof) {
  double tot = 0.0, tile = 0.0, fin = 0.0;
  for (int it = 0; it < iters; ++it) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, ctx->ev[2 * it], ctx->ev[2 * it + 2]) == hipSuccess) tot += ms;
    if (prof) {
      if (hipEventElapsedTime(&ms, ctx->ev[2 * it], ctx->ev[2 * it + 1]) == hipSuccess)
        tile += ms / kdh_ctx::kProfileReps;
      if (hipEventElapsedTime(&ms, ctx->ev[2 * it + 1], ctx->ev[2 * it + 2]) == hipSuccess) fin += ms;
    }
  }
  ctx->last_kernel_us = iters ? 1000.0 * tot / iters : 0.0;
  ctx->last_tile_us = (prof && iters) ? 1000.0 * tile / iters : 0.0;
  ctx->last_finish_us = (prof && iters) ? 1000.0 * fin / iters : 0.0;
  ctx->last_launches = iters;
}

hipGraphExec_t kdh_graph_lookup(kdh_ctx::GraphSlot& slot, const long long* key) {
  if (slot.exec && memcmp(slot.key, key, sizeof(slot.key)) == 0) return slot.exec;
  if (slot.exec) {
    (void)hipGraphExecDestroy(slot.exec);
    slot.exec = nullptr;
  }
  return nullptr;
}

void kdh_graph_store(kdh_ctx::GraphSlot& slot, const long long* key, hipGraphExec_t exec) {
  memcpy(slot.key, key, sizeof(slot.key));
  slot.exec = exec;
}

int kdh_last_kernel_time(kdh_ctx* ctx, double* avg_us, int* n_launches) {
  if (!ctx) return KDH_ERR_INVALID;
  if (avg_us) *avg_us = ctx->last_kernel_us;
  if (n_launches) *n_launches = ctx->last_launches;
  return KDH_OK;
}

int kdh_last_kernel_breakdown(kdh_ctx* ctx, double* tile_us, double* finish_us) {
  if (!ctx) return KDH_ERR_INVALID;
  if (tile_us) *tile_us = ctx->last_tile_us;
  if (finish_us) *finish_us = ctx->last_finish_us;
  return KDH_OK;
}

int kdh_set_profiling(kdh_ctx* ctx, int on) {
  if (!ctx) return KDH_ERR_INVALID;
  std::lock_guard<std::mutex> lk(ctx->mu);
  ctx->profile = on ? 1 : 0;
  return KDH_OK;
}
