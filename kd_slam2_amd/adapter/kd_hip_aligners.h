// ICPHIP3D / CTICPHIP3D: the configurable aligner types of the MI355X path,
// with the interface of the reference's TreeAlignerBase_ / ICP_ / CTICP_
// (kd_slam2/src/lib/kd_slam/icp/tree_aligner_base_.h:26-161,
//  icp/icp_.h:53-161, cticp/cticp_.h:20-261) over the C ABI of libkdhip.so.
//
// A .conf pipeline selects them by type name exactly as it selects
// ICPCPU3D / ICPCUDA3D (cfg/make_cpu_confs.sh:13-18); the registration and
// the tree-conversion hook follow the reference's plugin pattern
// (d3/instances_cuda.cpp:7-11, d3/tree.cu:22-30).
//
// Two execution forms:
//   compute()          the plain per-call path: a host Gauss-Newton loop that
//                      calls kdh_icp_build / kdh_cticp_build once per
//                      iteration (the reference's _buildQuadraticForm
//                      override point, icp/icp_.h:149) and solves the damped
//                      system in FP64 on the host;
//   computeOnDevice()  the whole loop resident on the GPU (kdh_align).
#pragma once

#include <array>
#include <memory>
#include <map>
#include <string>
#include <vector>

#include "../../include/kdh.h"
#include "../../include/kdh_host.h"
#include "srrg_standin.h"

namespace kd_slam {
namespace hip {

// 3x4 row-major rigid transform (the adapter's stand-in for Eigen::Isometry3f).
struct Isometry3f {
  std::array<float, 12> m{1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
  static Isometry3f Identity() { return Isometry3f(); }
};
using Vector6f = std::array<float, 6>;
using Matrix6f = std::array<float, 36>;

// Process-wide HIP context per device (the reference sets its device flags
// once at library load, kd_slam_cuda_init.cu:3-6).
kdh_ctx* defaultContext(int device = 0);

// ---- trees --------------------------------------------------------------
// Host tree as produced by TreeGenerator3D (Node_ array + leaf list).
struct TreeCPU3D {
  std::vector<kdh_node3f> nodes;
  std::vector<uint64_t> leaves;
  bool isGPU() const { return false; }
};

// Tree resident in HBM (TreeCUDA_ counterpart, tree/tree_cuda_.h:14-62).
struct TreeHIP3D {
  TreeHIP3D(kdh_ctx* ctx, const kdh_node3f* nodes, int64_t n_nodes, const uint64_t* leaves, int64_t n_leaves);
  ~TreeHIP3D();
  TreeHIP3D(const TreeHIP3D&) = delete;
  TreeHIP3D& operator=(const TreeHIP3D&) = delete;
  bool isGPU() const { return true; }
  kdh_tree* handle() const { return h_; }
  int64_t numLeaves() const { return n_leaves_; }

 private:
  kdh_ctx* ctx_;
  kdh_tree* h_ = nullptr;
  int64_t n_leaves_ = 0;
};

// Tree_toCompute (tree/tree_convert_.h:11-28): the function pointer a GPU
// plugin replaces at load time; with this library loaded it uploads.
using ToComputeFn = std::shared_ptr<TreeHIP3D> (*)(const TreeCPU3D&);
extern ToComputeFn g_tree_toCompute3d;

// ---- statistics and termination (icp/icp_stats_.h) ----------------------
struct ICPStats {
  int num_inliers = 0, num_outliers = 0, num_bad = 0, num_err_normal = 0, num_err_fn = 0;
  float chi2 = 0, chi2_kernelized = 0, pert_pose_norm = 0, pert_vel_norm = 0;
  int numLeaves() const { return num_inliers + num_outliers + num_bad + num_err_normal + num_err_fn; }
  float inlierRatio() const { int n = numLeaves(); return n ? (float)num_inliers / (float)n : 0.f; }
  float score() const { return inlierRatio() / (chi2_kernelized / num_inliers + 1e-6f); }
};

struct ICPTerminationCriteriaScore3D : public srrg2_core::Configurable {
  PARAM(srrg2_core::PropertyFloat, d_pert, "perturbation change, smaller than that terminates", 1e-3f, nullptr);
  PARAM(srrg2_core::PropertyFloat, d_score, "score change between iterations, smaller than that terminates", 1e-3f,
        nullptr);
  float prev_score = 0;
  bool hasToStop(ICPStats& s, int it);
};

enum AlignerType { AlignerICP, AlignerCTICP };

// TreeAlignerBase_::FixedEntry_ (icp/tree_aligner_base_.h:57-74): one tree of
// the fixed forest at its own pose, and its share of the last round.
template <int D>
struct FixedEntryHIP {
  const kdh_tree* fixed_tree = nullptr;
  double fixed_X[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
  ICPStats stats;
  double H[D * D] = {};
  double b[D] = {};
};

// ---- the aligners ---------------------------------------------------------
template <int CT>
struct AlignerHIP3D : public srrg2_core::Configurable {
  static constexpr int D = CT ? 12 : 6;
  PARAM(srrg2_core::PropertyFloat, kernel_threshold, "chi2 kernel threshold", 0.1f, &_params_changed);
  PARAM(srrg2_core::PropertyFloat, damping, "linear system damping", 1.0f, &_params_changed);
  PARAM(srrg2_core::PropertyFloat, max_error, "max point-to-plane error", 1.0f, &_params_changed);
  PARAM(srrg2_core::PropertyFloat, min_normal_cos, "min cosine of normal angle", 0.906307787f, &_params_changed);
  PARAM(srrg2_core::PropertyInt, min_good_measurements, "min inliers to accept a round", 10, &_params_changed);
  PARAM(srrg2_core::PropertyInt, max_iterations, "max ICP iterations", 20, &_params_changed);
  PARAM(srrg2_core::PropertyString, log_file, "file where to dump aligner log", "", &_params_changed);
  PARAM(srrg2_core::PropertyConfigurable_<ICPTerminationCriteriaScore3D>, term_crit, "termination criteria",
        nullptr, &_params_changed);
  // CTICP_ only (cticp/cticp_.h:61); ignored by the rigid aligner.
  PARAM(srrg2_core::PropertyFloat, velocity_damping, "velocity regularization weight", 100.f, &_params_changed);
  // Jacobian re-centring about the scan centroid (FP32 terms stay small far
  // from the origin); 0 reproduces the reference formula.
  PARAM(srrg2_core::PropertyBool, recenter, "re-centre the pose Jacobian on the scan centroid", true,
        &_params_changed);

  AlignerHIP3D();
  virtual ~AlignerHIP3D() = default;

  void setContext(kdh_ctx* ctx) { _ctx = ctx; }
  void setFixed(const TreeHIP3D& fixed) { _fixed = fixed.handle(); }
  void setMoving(const TreeHIP3D& moving) { _moving = moving.handle(); }
  // Trees already uploaded through the C ABI.
  void setFixed(const kdh_tree* fixed) { _fixed = fixed; }
  void setMoving(const kdh_tree* moving) { _moving = moving; }
  AlignerType alignerType() const { return CT ? AlignerCTICP : AlignerICP; }
  bool isGPU() const { return true; }

  Isometry3f getX() const;
  void setX(const Isometry3f& X);
  void setFixedX(const Isometry3f& X);
  Vector6f velocities() const;
  void setVelocities(const Vector6f& v);
  void resetMovingState();
  void setPosePrior(const Isometry3f& X_ref, const Isometry3f& Z, const Matrix6f& info);

  // ---- fixed forest (icp/icp_.h:90-95, icp/icp_impl_.h:30-71) ----
  using FixedEntry = FixedEntryHIP<D>;
  std::shared_ptr<FixedEntry> makeFixedEntry(const TreeHIP3D& tree, const Isometry3f& X);
  std::shared_ptr<FixedEntry> makeFixedEntry(const kdh_tree* tree, const double X[12]);
  // std::map: entries are summed in key order, as the reference iterates it
  std::map<int, std::shared_ptr<FixedEntry>> _fixed_forest;

  Matrix6f getPoseHessian() const;
  void saveStartState();
  std::string iterationLogHeader() const;
  void dumpIterationLog();

  void syncParams();
  void buildQuadraticForm(bool stats_mode = false);
  void buildQuadraticForm(FixedEntry& entry, bool stats_mode = false);
  bool oneRound(bool stats_mode = false);
  int compute(bool stats_mode = false);
  int computeOnDevice(bool stats_mode = false);

  // FP64 state (the drop-in keeps Eigen FP32 types at its surface)
  double X[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
  double Xf[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
  double v[6] = {0, 0, 0, 0, 0, 0};
  double H[D * D] = {};
  double b[D] = {};
  double dx[D] = {};
  ICPStats stats;
  std::vector<ICPStats> stats_log;
  kdh_params params{};

 protected:
  bool _params_changed = true;
  kdh_ctx* _ctx = nullptr;
  const kdh_tree* _fixed = nullptr;
  const kdh_tree* _moving = nullptr;
  bool _has_prior = false;
  kdh_prior _prior{};
  double _start_X[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
  double _start_Xf[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
  double _start_v[6] = {0, 0, 0, 0, 0, 0};
  std::unique_ptr<std::ostream> _log_stream;
  std::string _log_path;
};

struct ICPHIP3D : public AlignerHIP3D<0> {};
struct CTICPHIP3D : public AlignerHIP3D<1> {
  // fixed_state.velocities (cticp/cticp_.h): the fixed cloud's motion, used
  // only by the dual form
  void setFixedVelocities(const Vector6f& v);
  // CTICP_::buildQuadraticFormDual (cticp/cticp_impl_.h:42-45): H, b and the
  // fixed side's velocity blocks (cticp/cticp_.h:179-186), row-major
  void buildQuadraticFormDual(bool stats_mode = false);
  double vf[6] = {0, 0, 0, 0, 0, 0};
  double H_dual_vel_B[36] = {};
  double b_dual_vel_B[6] = {};
  double H_dual_cross_pose_vel_B[36] = {};  // rows: pose
  double H_dual_cross_vel_AB[36] = {};      // rows: moving velocities
};

void registerHIPTypes3D();

}  // namespace hip
}  // namespace kd_slam
