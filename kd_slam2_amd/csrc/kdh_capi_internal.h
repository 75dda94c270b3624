// Opaque C-ABI handles and helpers shared by the host-side translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/kdh.h"
#include "kdh_internal.h"

struct kdh_tree {
  int device = 0;
  void* dmem = nullptr;
  size_t bytes = 0;
  kdh::DevTree dt{};
  kdh::V3f root_mean{};
  float* node_dts = nullptr;  // per node (float)_dts_mean, for the in-place deskew
  // trees built on the device (kdh_tree_build) also keep, in dmem, the full
  // Node_ records, the leaf list and the permuted points
  kdh_node3f* full_nodes = nullptr;
  uint64_t* full_leaves = nullptr;
  float* full_xyz = nullptr;
  double* full_stamps = nullptr;
  int64_t n_points = 0;
};

struct kdh_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::mutex mu;
  std::string err;
  // workspace (grown x2 on demand, never shrunk)
  kdh::PairState* d_pairs = nullptr;
  size_t pairs_cap = 0;
  kdh::PairState* h_pairs = nullptr;  // pinned mirror
  kdh::PairState* h_pairs_dev = nullptr;  // its device address (kernels write results into it)
  size_t h_pairs_cap = 0;
  float* d_partials = nullptr;
  size_t partials_cap = 0;
  kdh::TileDesc* d_tiles = nullptr;
  size_t tile_cap = 0;
  // pinned staging of the tile list, and the list the device holds now (an
  // unchanged batch skips the upload)
  kdh::TileDesc* h_tiles = nullptr;
  size_t h_tile_cap = 0;
  std::vector<kdh::TileDesc> tiles_on_device;
  int* d_corr = nullptr;
  size_t corr_cap = 0;
  int* d_depth = nullptr;
  size_t depth_cap = 0;
  float* d_extract = nullptr;  // staging of tree read-backs
  float* h_root = nullptr;      // pinned, mapped: the root mean an in-place tree op writes back
  float* h_root_dev = nullptr;
  size_t extract_cap = 0;
  void* tree_build_ws = nullptr;  // device tree builder workspace (kdh_tree_build.hip)
  // Instantiated graphs, reused while the launch shape and buffers match.
  struct GraphSlot {
    hipGraphExec_t exec = nullptr;
    long long key[16] = {};
  };
  GraphSlot graph_align, graph_batch;
  // launch configuration (autotuner)
  int grid_override = 0;
  int lds_nodes = 256;  // measured: deeper staging costs more per tile than its LDS hits save
  int leaves_per_lane = 1;  // moving leaves per lane of the tile kernel: 1 or 2
  int tile_leaves = kdh::kTile;  // moving leaves per tile: 256, or 128 (one leaf per lane)
  int batch_streams = 1;    // kdh_batch_align replay on 1 or 2 streams
  // second stream (default priority) and fork / stagger / join events of the
  // two-stream batch replay, made on first use
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_stag = nullptr, ev_join = nullptr;
  int xcd_remap = 1;
  // timing of the last device loop
  static constexpr int kEvents = 2 * kdh::kMaxIter + 2;
  static constexpr int kProfileReps = 4;
  hipEvent_t ev[kEvents] = {};
  double last_kernel_us = 0.0;   // tile + finish per iteration
  double last_tile_us = 0.0;     // tile kernel alone (profiling mode)
  double last_finish_us = 0.0;   // finish kernel alone (profiling mode)
  int last_launches = 0;
  int profile = 0;               // 1: no graph, events around each kernel
};

int kdh_fail(kdh_ctx* ctx, int code, const std::string& msg);
void kdh_tree_build_ws_free(kdh_ctx* ctx);
int kdh_hip_fail(kdh_ctx* ctx, hipError_t e, const char* where);
int kdh_ensure_pairs(kdh_ctx* ctx, size_t n_pairs);
int kdh_ensure_partials(kdh_ctx* ctx, size_t n_floats);
// Enqueues the copy of the evolving fields (X, v, iter, status, prev_score)
// and the statistics logs of pairs [0, n) from the device states into the
// pinned mirror; the caller synchronizes ctx->stream.
int kdh_pull_results(kdh_ctx* ctx, int n_pairs);
// H, b (and the dual blocks when dual), stats_log[0] of build mode, the same way.
int kdh_pull_forms(kdh_ctx* ctx, int n_pairs, int dual);
int kdh_ensure_tiles(kdh_ctx* ctx, size_t n);
// Uploads the tile list (through pinned staging) unless the device already
// holds exactly this list.
int kdh_upload_tiles(kdh_ctx* ctx, const std::vector<kdh::TileDesc>& tiles);
// Two tile lists hold the same records.
bool kdh_same_tiles(const std::vector<kdh::TileDesc>& a, const std::vector<kdh::TileDesc>& b);
int kdh_fill_pair(kdh_ctx* ctx, kdh::PairState& ps, const kdh_tree* fixed, const kdh_tree* moving, int ct,
                  const kdh_params& p, int stats_mode, int mode);
void kdh_cache_from_state(kdh::PairState& ps);
int kdh_setup_prior(kdh::PairState& ps, const kdh_prior* prior);
int kdh_launch(kdh_ctx* ctx, kdh::PairState* d_pairs, const kdh::TileDesc* d_tiles, int n_tiles, int ct, int* corr,
               int* depth, float* partials, int n_pairs, int phases = 3, int dual = 0, int pair_first = 0,
               hipStream_t stream = nullptr);
// compute()'s return value from a finished state: -i after a failed
// min_good gate in round i, i after a stop by the criterion, else
// max_iterations.
int kdh_result_code(const kdh::PairState& ps, int max_iterations_param);
int kdh_result_code_of(int status, int iter, int max_iterations_param);
// Workgroups of a tile launch over n_tiles tiles: the context's persistent
// grid, or one per tile; with the XCD remap on, rounded up to a multiple of 8
// so that every XCD takes a contiguous range of tiles (the extra workgroups of
// the last range find no tile).
inline int kdh_tile_grid(const kdh_ctx* ctx, int n_tiles) {
  int g = ctx->grid_override > 0 ? (ctx->grid_override < n_tiles ? ctx->grid_override : n_tiles) : n_tiles;
  if (ctx->xcd_remap && g >= 8) g = (g + 7) & ~7;
  return g;
}
void kdh_collect_times(kdh_ctx* ctx, int iters, bool prof);
// Returns the cached executable graph for `key`, or nullptr (and drops the
// stale one) when the key differs.
hipGraphExec_t kdh_graph_lookup(kdh_ctx::GraphSlot& slot, const long long* key);
void kdh_graph_store(kdh_ctx::GraphSlot& slot, const long long* key, hipGraphExec_t exec);
