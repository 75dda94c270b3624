// Launch interface of the tile kernel (kdh_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "kdh_internal.h"

namespace kdh {

struct LaunchArgs {
  PairState* pairs;       // device array of alignments
  const TileDesc* tiles;  // device array, one record per tile (nullptr: all tiles belong to pair 0)
  int n_tiles_total;
  float* partials;        // n_tiles_total x slots(ct, dual) floats
  int* corr_out;          // optional, single-pair launches: fixed leaf per moving leaf
  int* depth_out;         // optional, single-pair launches: descent depth per moving leaf
  int grid;               // workgroups (tiles are strided over them)
  int ct;
  int lds_nodes;          // top-of-tree nodes staged in LDS: 0, 256, 512, 1024, 2048 (ICP)
  int xcd_remap;          // 1: give each XCD a contiguous range of tiles (needs grid % 8 == 0)
  int pair_first;         // finish stage: alignments [pair_first, pair_first + n_pairs)
  int n_pairs;            // (0: no finish launch)
  int phases;             // bit 0: tile kernel, bit 1: finish stage
  const int* pair_tile_first;  // optional, finish stage: first tile of each pair in this launch (else PairState::tile_first)
  int* status_out;        // optional: per-pair status mirror written by the finish stage
  int dual = 0;           // 1: CT dual form of a bundle-adjustment factor (build mode, ct = 1)
  int leaves_per_lane = 1;  // 1, or 2: two descents in flight per lane (256-leaf tiles, not the dual form)
  int tile_leaves = 256;    // moving leaves per tile: 256, or 128 (one leaf per lane)
};

hipError_t launch_tiles(const LaunchArgs& a, hipStream_t s);
int slots(int ct, int dual = 0);

}  // namespace kdh
