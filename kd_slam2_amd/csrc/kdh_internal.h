// Device-side data structures shared by the kernels and the C-ABI layer.
//
// HBM layout of a tree (one allocation per tree, 256-byte aligned sections):
//   nodes  : 2 x float4 per node, BFS order as produced by the reference's
//            prune + breadthFirstPermutation (tree_cpu_.h:173-207,
//            tree_cpu_impl_.h:137-158):
//              a = (mean.x, mean.y, mean.z, dir.x)
//              b = (dir.y, dir.z, left, right)      (links as int32 bits)
//            32 bytes per node, one 32-byte sector per visit.
//   leaves : 2 x float4 per leaf of _leaves_indices order (the query set when
//            the tree is the moving cloud):
//              a = (mean.x, mean.y, mean.z, dir.x)
//              b = (dir.y, dir.z, (float)_dts_mean, node index bits)
// The 64-byte host node (Node_<Point3fTraits>, node_.h:18-43) also carries
// _idx_start/_idx_end and _dts_covariance, which the alignment never reads.
#pragma once

#include <stddef.h>

#include "kdh_math.h"
#include "kdh_term.h"

namespace kdh {

constexpr int kTile = 256;           // moving leaves per tile (one workgroup pass)
constexpr int kMaxIter = 64;         // longest Gauss-Newton loop kept on device

struct DevTree {
  const float4* nodes;
  const float4* leaves;
  int n_nodes;
  int n_leaves;
  const float* node_dts;  // per node (float)_dts_mean
};

// Per-iteration statistics, ICPStats_ (icp_stats_.h:10-20) in FP64 sums.
struct DevStats {
  int num_inliers, num_outliers, num_bad, num_err_normal, num_err_fn;
  int pad;
  double chi2, chi2_kernelized, pert_pose_norm, pert_vel_norm;
};

// Solver settings of one alignment (ICPParams_ / CTICPParams_,
// tree_aligner_base_.h:12-21, cticp_.h:14-18, and the score termination
// criterion icp_stats_.h:85-102).
struct SolveParams {
  double damping;
  double velocity_damping;
  int min_good_measurements;
  int max_iterations;
  int use_term_crit;
  float d_pert;
  float d_score;
  int recenter;
};

// Everything the device needs to run one alignment, and its evolving state.
struct PairState {
  // inputs
  DevTree fixed;
  DevTree moving;
  TermParams tp;
  SolveParams sp;
  int ct;               // 0 rigid ICP, 1 CT-ICP
  int tile_first;       // first tile of this pair in the launch's tile list
  int n_tiles;
  int mode;             // 0: solve + update (align), 1: build only (write H/b)
  // Fixed forest (_fixed_forest, icp_impl_.h:53-71): a group of consecutive
  // pair states with one moving tree and pose, one fixed tree and pose each.
  // The first (leader) holds group_n = entries and group_tiles = their tiles
  // (contiguous in the launch); the others hold group_n = 0.  A plain pair
  // is a group of one.
  int group_n;
  int group_tiles;
  int group_pad[4];     // keeps sizeof(PairState) a multiple of 16
  double Xf[12];        // fixed pose, 3x4 row-major
  double has_prior;
  double prior_ZXi[12]; // (X_ref * Z)^-1
  double prior_info[36];
  V3f moving_root_mean; // re-centring anchor in the moving frame
  double c64[3];        // the anchor in world coordinates, FP64 (0 when not re-centred)
  // state
  double X[12];         // moving pose
  double v[6];          // moving velocities (CT)
  TermCache cache;      // FP32 inputs of the current iteration
  int iter;             // completed iterations
  int status;           // 0 running, 1 finished (max_iterations), 2 failed min_good gate, 3 stopped by criterion
  float prev_score;
  // outputs
  DevStats stats_log[kMaxIter];
  double H[144];        // last H (n x n row-major, n = 6 or 12)
  double b[12];
  // dual form (build mode, kdh_batch_build_dual): H_vel_B 36, b_vel_B 6,
  // H_cross_pose_vel_B 36, H_cross_vel_AB 36
  double Hd[114];
};

// Bytes of a PairState the host fills (inputs and evolving state, up to the
// statistics log); the rest is written by the device before it is read.
constexpr size_t kStateBytes = (offsetof(PairState, stats_log) + 15) & ~(size_t)15;

// One tile of a launch: its alignment and what the tile kernel needs of that
// alignment before it can load the tile's leaves and stage the top of the
// fixed tree, in one 32-byte record (one scalar load on the device; from the
// pair state they would take a second round trip after the tile's pair).
struct alignas(32) TileDesc {
  const float4* nodes;   // fixed tree records
  const float4* leaves;  // moving tree leaves
  int n_nodes;
  int n_leaves;
  int pair;              // index of the pair state
  int leaf0;             // first moving leaf of the tile
};
static_assert(sizeof(TileDesc) == 32, "a tile record is one 32-byte load");

// Tile `tile` (0-based within its alignment) of pair state `pair`, with
// `tile_leaves` moving leaves per tile.
KDH_HD TileDesc make_tile_desc(const PairState& ps, int pair, int tile, int tile_leaves) {
  TileDesc d;
  d.nodes = ps.fixed.nodes;
  d.leaves = ps.moving.leaves;
  d.n_nodes = ps.fixed.n_nodes;
  d.n_leaves = ps.moving.n_leaves;
  d.pair = pair;
  d.leaf0 = tile * tile_leaves;
  return d;
}

// Re-centring anchor of a state (the TermCache fields mroot and cf, and the
// FP64 anchor the solver maps the form back with): c = Xm * mroot in FP64,
// cf = Xf^-1 * c rounded once.  All zero when re-centring is off.
struct Anchor {
  V3f mroot, cf;
  V3d c;
};
KDH_HD Anchor make_anchor(const IsoD& Xm, const IsoD& Xf, V3f mroot, int on) {
  Anchor a;
  if (!on) {
    a.mroot = a.cf = v3(0.0f, 0.0f, 0.0f);
    a.c = v3(0.0, 0.0, 0.0);
    return a;
  }
  a.c = apply_iso(Xm, v3((double)mroot.x, (double)mroot.y, (double)mroot.z));
  const V3d cf = apply_iso(inverse(Xf), a.c);
  a.mroot = mroot;
  a.cf = v3((float)cf.x, (float)cf.y, (float)cf.z);
  return a;
}
// Writes the anchor into a cache and the state's FP64 copy.
KDH_HD void set_anchor(TermCache& C, double* c64, const Anchor& a) {
  C.mroot = a.mroot;
  C.cf = a.cf;
  c64[0] = a.c.x;
  c64[1] = a.c.y;
  c64[2] = a.c.z;
}

// The FP32 cache of a state's FP64 pose, fixed pose and velocities, with its
// re-centring anchor (the host set-up and the device's batch initialisation
// evaluate the same expressions, so the cache bits agree).
KDH_HD void cache_from_pose(TermCache& C, double c64[3], const double* X, const double* Xf34, const double* v,
                            V3f mroot, int recenter) {
  const IsoD Xc = iso_from_34<double>(X);
  const IsoD Xf = iso_from_34<double>(Xf34);
  C.Xm = iso_cast<float>(Xc);
  C.Xf = iso_cast<float>(Xf);
  C.Xmf = iso_cast<float>(compose(inverse(Xf), Xc));
  for (int i = 0; i < 6; ++i) C.v[i] = (float)v[i];
  set_anchor(C, c64, make_anchor(Xc, Xf, mroot, recenter));
}
KDH_HD void cache_from_state(PairState& ps) {
  cache_from_pose(ps.cache, ps.c64, ps.X, ps.Xf, ps.v, ps.moving_root_mean, ps.sp.recenter);
}

}  // namespace kdh
