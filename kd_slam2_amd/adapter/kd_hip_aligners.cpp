// Implementation of the ICPHIP3D / CTICPHIP3D adapter (see kd_hip_aligners.h).
#include "kd_hip_aligners.h"

#include <math.h>
#include <string.h>

#include <fstream>
#include <mutex>
#include <sstream>
#include <stdexcept>

#include "../csrc/kdh_solve.h"

namespace kd_slam {
namespace hip {

namespace {
std::mutex g_ctx_mu;
kdh_ctx* g_ctx[16] = {};

void check(kdh_ctx* ctx, int rc, const char* what) {
  if (rc != KDH_OK) throw std::runtime_error(std::string(what) + ": " + kdh_last_error(ctx));
}

// updateCache (icp/icp_impl_.h:16-21): FP32 Xm, Xf, Xmf = Xf^-1 Xm, formed
// from the FP64 state and rounded once (the device loop does the same).
void make_cache(const double* X, const double* Xf, float* Xm32, float* Xf32, float* Xmf32) {
  kdh::IsoD m = kdh::iso_from_34<double>(X);
  kdh::IsoD f = kdh::iso_from_34<double>(Xf);
  kdh::iso_to_34(m, Xm32);
  kdh::iso_to_34(f, Xf32);
  kdh::iso_to_34(kdh::compose(kdh::inverse(f), m), Xmf32);
}
}  // namespace

kdh_ctx* defaultContext(int device) {
  if (device < 0 || device >= 16) throw std::runtime_error("device out of range");
  std::lock_guard<std::mutex> lk(g_ctx_mu);
  if (!g_ctx[device]) {
    int rc = kdh_ctx_create(device, &g_ctx[device]);
    if (rc != KDH_OK) throw std::runtime_error("kdh_ctx_create failed");
  }
  return g_ctx[device];
}

TreeHIP3D::TreeHIP3D(kdh_ctx* ctx, const kdh_node3f* nodes, int64_t n_nodes, const uint64_t* leaves,
                     int64_t n_leaves)
    : ctx_(ctx), n_leaves_(n_leaves) {
  check(ctx_, kdh_tree_upload(ctx_, nodes, n_nodes, leaves, n_leaves, &h_), "kdh_tree_upload");
}

TreeHIP3D::~TreeHIP3D() {
  if (h_) kdh_tree_free(ctx_, h_);
}

static std::shared_ptr<TreeHIP3D> toComputeCPUOnly(const TreeCPU3D&) {
  throw std::runtime_error("no GPU library loaded");
}
static std::shared_ptr<TreeHIP3D> toComputeHIP(const TreeCPU3D& t) {
  return std::make_shared<TreeHIP3D>(defaultContext(0), t.nodes.data(), (int64_t)t.nodes.size(), t.leaves.data(),
                                     (int64_t)t.leaves.size());
}
ToComputeFn g_tree_toCompute3d = toComputeCPUOnly;

bool ICPTerminationCriteriaScore3D::hasToStop(ICPStats& s, int it) {
  if (it > 0) {
    bool pert_ok = s.pert_pose_norm < param_d_pert.value();
    float ds = s.score() - prev_score;
    bool score_ok = (ds > 0) && ds / (prev_score + 1e-6) < param_d_score.value();
    if (pert_ok && score_ok) return true;
  }
  prev_score = s.score();
  return false;
}

template <int CT>
AlignerHIP3D<CT>::AlignerHIP3D() {
  kdh_params_default(&params);
}

template <int CT>
Isometry3f AlignerHIP3D<CT>::getX() const {
  Isometry3f r;
  for (int i = 0; i < 12; ++i) r.m[i] = (float)X[i];
  return r;
}
template <int CT>
void AlignerHIP3D<CT>::setX(const Isometry3f& x) {
  for (int i = 0; i < 12; ++i) X[i] = x.m[i];
}
template <int CT>
void AlignerHIP3D<CT>::setFixedX(const Isometry3f& x) {
  for (int i = 0; i < 12; ++i) Xf[i] = x.m[i];
  if (CT) {
    // CTICP_::setFixedX also clears the fixed velocities (cticp_.h:222-225);
    // the fixed cloud carries no velocity state here.
  }
}
template <int CT>
Vector6f AlignerHIP3D<CT>::velocities() const {
  Vector6f r{};
  for (int i = 0; i < 6; ++i) r[i] = CT ? (float)v[i] : 0.f;
  return r;
}
template <int CT>
void AlignerHIP3D<CT>::setVelocities(const Vector6f& vv) {
  if (!CT) return;  // TreeAlignerBase_ default: no-op
  for (int i = 0; i < 6; ++i) v[i] = vv[i];
}
template <int CT>
void AlignerHIP3D<CT>::resetMovingState() {
  for (int i = 0; i < 6; ++i) v[i] = 0;
}
template <int CT>
void AlignerHIP3D<CT>::setPosePrior(const Isometry3f& X_ref, const Isometry3f& Z, const Matrix6f& info) {
  _has_prior = false;
  for (int i = 0; i < 36; ++i) _has_prior = _has_prior || info[i] != 0.f;
  for (int i = 0; i < 12; ++i) {
    _prior.X_ref[i] = X_ref.m[i];
    _prior.Z[i] = Z.m[i];
  }
  for (int i = 0; i < 36; ++i) _prior.info[i] = info[i];
}

template <int CT>
void AlignerHIP3D<CT>::syncParams() {
  if (!_params_changed) return;
  params.kernel_threshold = param_kernel_threshold.value();
  params.damping = param_damping.value();
  params.max_error = param_max_error.value();
  params.min_normal_cos = param_min_normal_cos.value();
  params.min_good_measurements = param_min_good_measurements.value();
  params.max_iterations = param_max_iterations.value();
  params.velocity_damping = param_velocity_damping.value();
  params.recenter = param_recenter.value() ? 1 : 0;
  auto tc = param_term_crit.value();
  params.use_term_crit = tc ? 1 : 0;
  if (tc) {
    params.term_d_pert = tc->param_d_pert.value();
    params.term_d_score = tc->param_d_score.value();
  }
  // log_file (TreeAlignerBase_::syncParams, tree_aligner_base_.h:77-84)
  const std::string& lf = param_log_file.value();
  if (lf.empty()) {
    _log_stream.reset();
    _log_path.clear();
  } else if (!_log_stream || lf != _log_path) {
    _log_stream.reset(new std::ofstream(lf));
    _log_path = lf;
  }
  _params_changed = false;
}

// getPoseHessian (icp/icp_.h:47, cticp/cticp_.h:169-177, 226): the pose block
// of H; for CT-ICP the velocities are marginalised out by the Schur
// complement H_pp - H_pv (H_vv + 1e-9 I)^-1 H_pv^T (FP64 here).
template <int CT>
Matrix6f AlignerHIP3D<CT>::getPoseHessian() const {
  Matrix6f out{};
  double P[36];
  for (int r = 0; r < 6; ++r)
    for (int c = 0; c < 6; ++c) P[r * 6 + c] = H[r * D + c];
  if (CT) {
    // columns of (H_vv + 1e-9 I)^-1 H_pv^T, one LDL^T solve per pose row
    double Hvv[36];
    for (int r = 0; r < 6; ++r)
      for (int c = 0; c < 6; ++c) Hvv[r * 6 + c] = H[(6 + r) * D + 6 + c] + (r == c ? 1e-9 : 0.0);
    for (int j = 0; j < 6; ++j) {
      double rhs[6], y[6];
      for (int k = 0; k < 6; ++k) rhs[k] = H[j * D + 6 + k];  // row j of H_pv
      kdh::ldlt_solve_d<6>(Hvv, rhs, y);
      for (int i = 0; i < 6; ++i) {
        double s = 0;
        for (int k = 0; k < 6; ++k) s += H[i * D + 6 + k] * y[k];
        P[i * 6 + j] -= s;
      }
    }
  }
  for (int i = 0; i < 36; ++i) out[i] = (float)P[i];
  return out;
}

// saveStartState (icp/icp_.h:142-145, cticp/cticp_.h:248-251)
template <int CT>
void AlignerHIP3D<CT>::saveStartState() {
  memcpy(_start_X, X, sizeof(X));
  memcpy(_start_Xf, Xf, sizeof(Xf));
  memcpy(_start_v, v, sizeof(v));
}

// iterationLogHeader (icp/icp_impl_.h:155-169, cticp/cticp_impl_.h:296-310):
// fixed/moving poses as logmap and velocities, at the start and now.
template <int CT>
std::string AlignerHIP3D<CT>::iterationLogHeader() const {
  auto vec = [](const double* x, int n) {
    std::ostringstream os;
    for (int i = 0; i < n; ++i) os << (i ? " " : "") << (float)x[i];
    return os.str();
  };
  auto lm = [&](const double* X34) {
    double e[6];
    kdh::logmap_d(kdh::iso_from_34<double>(X34), e);
    return vec(e, 6);
  };
  const double zero[6] = {0, 0, 0, 0, 0, 0};
  std::ostringstream os;
  os << (CT ? "CTICP" : "ICP") << 3 << "D| START "
     << " fp: " << lm(_start_Xf) << " fv: " << (CT ? vec(zero, 6) : "none")
     << " mp: " << lm(_start_X) << " mv: " << (CT ? vec(_start_v, 6) : "none") << " | END "
     << " fp: " << lm(Xf) << " fv: " << (CT ? vec(zero, 6) : "none")
     << " mp: " << lm(X) << " mv: " << (CT ? vec(v, 6) : "none");
  return os.str();
}

// dumpIterationLog (tree_aligner_base_.h:118-124)
template <int CT>
void AlignerHIP3D<CT>::dumpIterationLog() {
  if (!_log_stream) return;
  *_log_stream << iterationLogHeader() << std::endl;
  for (size_t i = 0; i < stats_log.size(); ++i) {
    const ICPStats& s = stats_log[i];
    *_log_stream << "  iter: " << i << " stats: k2: " << s.chi2 << " kk: " << s.chi2_kernelized
                 << " ni: " << s.num_inliers << " no: " << s.num_outliers << " nb: " << s.num_bad
                 << " nn: " << s.num_err_normal << " nf: " << s.num_err_fn << " pp: " << s.pert_pose_norm
                 << " pv: " << s.pert_vel_norm << std::endl;
  }
}

template <int CT>
void AlignerHIP3D<CT>::buildQuadraticForm(bool stats_mode) {
  syncParams();
  if (!_fixed || !_moving) throw std::runtime_error("setFixed/setMoving not called");
  kdh_ctx* ctx = _ctx ? _ctx : defaultContext(0);
  float Xm32[12], Xf32[12], Xmf32[12];
  make_cache(X, Xf, Xm32, Xf32, Xmf32);
  kdh_stats st;
  if (CT) {
    float v32[6];
    for (int i = 0; i < 6; ++i) v32[i] = (float)v[i];
    check(ctx, kdh_cticp_build(ctx, _fixed, _moving, Xm32, Xf32, Xmf32, v32, &params,
                               stats_mode ? 1 : 0, H, b, &st),
          "kdh_cticp_build");
  } else {
    check(ctx, kdh_icp_build(ctx, _fixed, _moving, Xm32, Xf32, Xmf32, &params,
                             stats_mode ? 1 : 0, H, b, &st),
          "kdh_icp_build");
  }
  stats.num_inliers = st.num_inliers;
  stats.num_outliers = st.num_outliers;
  stats.num_bad = st.num_bad;
  stats.num_err_normal = st.num_err_normal;
  stats.num_err_fn = st.num_err_fn;
  stats.chi2 = (float)st.chi2;
  stats.chi2_kernelized = (float)st.chi2_kernelized;
  stats.pert_pose_norm = 0;
  stats.pert_vel_norm = 0;
}

template <int CT>
std::shared_ptr<typename AlignerHIP3D<CT>::FixedEntry> AlignerHIP3D<CT>::makeFixedEntry(const TreeHIP3D& tree,
                                                                                        const Isometry3f& X) {
  double X34[12];
  for (int i = 0; i < 12; ++i) X34[i] = X.m[i];  // 3x4 row-major
  return makeFixedEntry(tree.handle(), X34);
}

template <int CT>
std::shared_ptr<typename AlignerHIP3D<CT>::FixedEntry> AlignerHIP3D<CT>::makeFixedEntry(const kdh_tree* tree,
                                                                                        const double X34[12]) {
  auto e = std::make_shared<FixedEntry>();
  e->fixed_tree = tree;
  memcpy(e->fixed_X, X34, sizeof(e->fixed_X));
  return e;
}

// ICP_::buildQuadraticForm(FixedEntry&) (icp/icp_impl_.h:30-40): the form
// against one entry of the fixed forest, kept in the entry.
template <int CT>
void AlignerHIP3D<CT>::buildQuadraticForm(FixedEntry& entry, bool stats_mode) {
  const kdh_tree* fixed = _fixed;
  double Xf_keep[12];
  memcpy(Xf_keep, Xf, sizeof(Xf));
  _fixed = entry.fixed_tree;
  memcpy(Xf, entry.fixed_X, sizeof(Xf));
  try {
    buildQuadraticForm(stats_mode);
  } catch (...) {
    _fixed = fixed;
    memcpy(Xf, Xf_keep, sizeof(Xf));
    throw;
  }
  _fixed = fixed;
  memcpy(Xf, Xf_keep, sizeof(Xf));
  entry.stats = stats;
  memcpy(entry.H, H, sizeof(H));
  memcpy(entry.b, b, sizeof(b));
}

// ICP_::oneRound / CTICP_::oneRound (icp/icp_impl_.h:53-90, cticp/cticp_impl_.h:48-92)
// with the damped system solved in FP64.
template <int CT>
bool AlignerHIP3D<CT>::oneRound(bool stats_mode) {
  if (_fixed_forest.empty()) {
    buildQuadraticForm(stats_mode);
  } else {
    double tH[D * D] = {}, tb[D] = {};
    ICPStats ts;
    for (auto& [ref, e] : _fixed_forest) {
      buildQuadraticForm(*e, stats_mode);
      for (int i = 0; i < D * D; ++i) tH[i] += e->H[i];
      for (int i = 0; i < D; ++i) tb[i] += e->b[i];
      ts.num_inliers += e->stats.num_inliers;
      ts.num_outliers += e->stats.num_outliers;
      ts.num_bad += e->stats.num_bad;
      ts.num_err_normal += e->stats.num_err_normal;
      ts.num_err_fn += e->stats.num_err_fn;
      ts.chi2 += e->stats.chi2;
      ts.chi2_kernelized += e->stats.chi2_kernelized;
    }
    memcpy(H, tH, sizeof(H));
    memcpy(b, tb, sizeof(b));
    stats = ts;
  }
  double Hs[D * D], bs[D];
  memcpy(Hs, H, sizeof(Hs));
  memcpy(bs, b, sizeof(bs));
  if (_has_prior) {
    kdh::IsoD Xc = kdh::iso_from_34<double>(X);
    kdh::IsoD ZXi = kdh::inverse(kdh::compose(kdh::iso_from_34<double>(_prior.X_ref),
                                              kdh::iso_from_34<double>(_prior.Z)));
    double e[6], J[36];
    kdh::logmap_d(kdh::compose(ZXi, Xc), e);
    kdh::dPoseToPoseError_d(ZXi, Xc, J);
    for (int r = 0; r < 6; ++r) {
      for (int c = 0; c < 6; ++c) {
        double s = 0;
        for (int k = 0; k < 6; ++k)
          for (int m = 0; m < 6; ++m) s += J[k * 6 + r] * _prior.info[k * 6 + m] * J[m * 6 + c];
        Hs[r * D + c] += s;
      }
      double sb = 0;
      for (int k = 0; k < 6; ++k)
        for (int m = 0; m < 6; ++m) sb += J[k * 6 + r] * _prior.info[k * 6 + m] * e[m];
      bs[r] += sb;
    }
  }
  memcpy(H, Hs, sizeof(Hs));
  memcpy(b, bs, sizeof(bs));
  if (stats.num_outliers + stats.num_inliers < params.min_good_measurements) return false;
  double A[D * D], rhs[D];
  memcpy(A, Hs, sizeof(A));
  for (int i = 0; i < D; ++i) {
    A[i * D + i] += (i < 6) ? params.damping : params.velocity_damping;
    rhs[i] = -bs[i];
  }
  kdh::ldlt_solve_d<D>(A, rhs, dx);
  double pn = 0, vn = 0;
  for (int i = 0; i < 6; ++i) pn += dx[i] * dx[i];
  for (int i = 6; i < D; ++i) vn += dx[i] * dx[i];
  stats.pert_pose_norm = (float)sqrt(pn);
  stats.pert_vel_norm = (float)sqrt(vn);
  if (stats_mode) return true;
  kdh::IsoD Xn = kdh::compose(kdh::expmap_d(dx), kdh::iso_from_34<double>(X));
  kdh::iso_to_34(Xn, X);
  if (CT)
    for (int i = 0; i < 6; ++i) v[i] += dx[6 + i];
  return true;
}

// TreeAlignerBase_::compute (tree_aligner_base_.h:100-115)
template <int CT>
int AlignerHIP3D<CT>::compute(bool stats_mode) {
  syncParams();
  stats_log.clear();
  saveStartState();
  const int max_iter = stats_mode ? 1 : param_max_iterations.value();
  for (int i = 0; i < max_iter; ++i) {
    bool ok = oneRound(stats_mode);
    stats_log.push_back(stats);
    if (!ok) return -i;
    auto tc = param_term_crit.value();
    if (tc && tc->hasToStop(stats, i)) return i;
  }
  return param_max_iterations.value();
}

template <int CT>
int AlignerHIP3D<CT>::computeOnDevice(bool stats_mode) {
  syncParams();
  saveStartState();
  if (!_fixed_forest.empty()) {
    // one round without update per entry: the per-call path
    if (stats_mode) return compute(true);
    if (!_moving) throw std::runtime_error("setMoving not called");
    kdh_ctx* ctx = _ctx ? _ctx : defaultContext(0);
    const int n = (int)_fixed_forest.size();
    std::vector<kdh_pair> pairs(n);
    std::vector<kdh_pair_result> res(n);
    int j = 0;
    for (auto& [ref, e] : _fixed_forest) {
      kdh_pair& pr = pairs[j];
      memset(&pr, 0, sizeof(pr));
      pr.fixed = e->fixed_tree;
      pr.moving = _moving;
      pr.ct = CT;
      pr.forest = j == 0 ? n - 1 : 0;
      memcpy(pr.X0, X, sizeof(pr.X0));
      memcpy(pr.v0, v, sizeof(pr.v0));
      memcpy(pr.Xf, e->fixed_X, sizeof(pr.Xf));
      pr.params = params;
      pr.prior = (j == 0 && _has_prior) ? &_prior : nullptr;
      ++j;
    }
    check(ctx, kdh_batch_align(ctx, n, pairs.data(), res.data()), "kdh_batch_align");
    auto to_stats = [](const kdh_stats& k) {
      ICPStats s;
      s.num_inliers = k.num_inliers;
      s.num_outliers = k.num_outliers;
      s.num_bad = k.num_bad;
      s.num_err_normal = k.num_err_normal;
      s.num_err_fn = k.num_err_fn;
      s.chi2 = (float)k.chi2;
      s.chi2_kernelized = (float)k.chi2_kernelized;
      s.pert_pose_norm = (float)k.pert_pose_norm;
      s.pert_vel_norm = (float)k.pert_vel_norm;
      return s;
    };
    j = 0;
    for (auto& [ref, e] : _fixed_forest) e->stats = to_stats(res[j++].entry);
    memcpy(X, res[0].X, sizeof(X));
    if (CT) memcpy(v, res[0].v, sizeof(v));
    stats = to_stats(res[0].last);
    stats_log.assign(1, stats);
    return res[0].result;
  }
  if (!_fixed || !_moving) throw std::runtime_error("setFixed/setMoving not called");
  kdh_ctx* ctx = _ctx ? _ctx : defaultContext(0);
  std::vector<kdh_stats> log(std::max(1, param_max_iterations.value()));
  int n = 0, res = 0;
  double Xo[12], vo[6];
  check(ctx, kdh_align(ctx, _fixed, _moving, CT, X, v, Xf, &params,
                       _has_prior ? &_prior : nullptr, stats_mode ? 1 : 0, Xo, vo, log.data(), &n, &res, nullptr),
        "kdh_align");
  memcpy(X, Xo, sizeof(X));
  if (CT) memcpy(v, vo, sizeof(v));
  stats_log.clear();
  for (int i = 0; i < n; ++i) {
    ICPStats s;
    s.num_inliers = log[i].num_inliers;
    s.num_outliers = log[i].num_outliers;
    s.num_bad = log[i].num_bad;
    s.num_err_normal = log[i].num_err_normal;
    s.num_err_fn = log[i].num_err_fn;
    s.chi2 = (float)log[i].chi2;
    s.chi2_kernelized = (float)log[i].chi2_kernelized;
    s.pert_pose_norm = (float)log[i].pert_pose_norm;
    s.pert_vel_norm = (float)log[i].pert_vel_norm;
    stats_log.push_back(s);
  }
  if (n) stats = stats_log.back();
  return res;
}

template struct AlignerHIP3D<0>;
template struct AlignerHIP3D<1>;

void CTICPHIP3D::setFixedVelocities(const Vector6f& vv) {
  for (int i = 0; i < 6; ++i) vf[i] = vv[i];
}

void CTICPHIP3D::buildQuadraticFormDual(bool stats_mode) {
  syncParams();
  if (!_fixed || !_moving) throw std::runtime_error("setFixed/setMoving not called");
  kdh_ctx* ctx = _ctx ? _ctx : defaultContext(0);
  kdh_dual_pair in;
  memset(&in, 0, sizeof(in));
  in.fixed = _fixed;
  in.moving = _moving;
  memcpy(in.X_moving, X, sizeof(in.X_moving));
  memcpy(in.v_moving, v, sizeof(in.v_moving));
  memcpy(in.X_fixed, Xf, sizeof(in.X_fixed));
  memcpy(in.v_fixed, vf, sizeof(in.v_fixed));
  in.params = params;
  kdh_dual_form out;
  check(ctx, kdh_batch_build_dual(ctx, 1, &in, stats_mode ? 1 : 0, &out), "kdh_batch_build_dual");
  memcpy(H, out.H, sizeof(H));
  memcpy(b, out.b, sizeof(b));
  memcpy(H_dual_vel_B, out.H_dual_vel_B, sizeof(H_dual_vel_B));
  memcpy(b_dual_vel_B, out.b_dual_vel_B, sizeof(b_dual_vel_B));
  memcpy(H_dual_cross_pose_vel_B, out.H_dual_cross_pose_vel_B, sizeof(H_dual_cross_pose_vel_B));
  memcpy(H_dual_cross_vel_AB, out.H_dual_cross_vel_AB, sizeof(H_dual_cross_vel_AB));
  const kdh_stats& st = out.stats;
  stats.num_inliers = st.num_inliers;
  stats.num_outliers = st.num_outliers;
  stats.num_bad = st.num_bad;
  stats.num_err_normal = st.num_err_normal;
  stats.num_err_fn = st.num_err_fn;
  stats.chi2 = (float)st.chi2;
  stats.chi2_kernelized = (float)st.chi2_kernelized;
  stats.pert_pose_norm = 0;
  stats.pert_vel_norm = 0;
}

// ---- registration (d3/icp.cu:13-17, d3/cticp.cu:14-18, d3/tree.cu:22-30) ----
void registerHIPTypes3D() {
  BOSS_REGISTER_CLASS(ICPHIP3D);
  BOSS_REGISTER_CLASS(CTICPHIP3D);
  BOSS_REGISTER_CLASS(ICPTerminationCriteriaScore3D);
  g_tree_toCompute3d = toComputeHIP;
}

void __attribute__((constructor)) initializeHIPTypes3D() { registerHIPTypes3D(); }

}  // namespace hip
}  // namespace kd_slam

// ---- C entry points used by tests and bench.py ------------------------------
using namespace kd_slam::hip;

namespace {

template <typename A>
int run_aligner(A* al, kdh_ctx* ctx, const kdh_tree* fixed, const kdh_tree* moving, const double* X0,
                const double* v0, const double* Xf, const kdh_params* p, int on_device, int stats_mode,
                double* X_out, double* v_out, int* result, int* n_iter, kdh_stats* last, kdh_stats* log) {
  al->setContext(ctx);
  al->param_kernel_threshold.setValue(p->kernel_threshold);
  al->param_damping.setValue(p->damping);
  al->param_max_error.setValue(p->max_error);
  al->param_min_normal_cos.setValue(p->min_normal_cos);
  al->param_min_good_measurements.setValue(p->min_good_measurements);
  al->param_max_iterations.setValue(p->max_iterations);
  al->param_velocity_damping.setValue(p->velocity_damping);
  al->param_recenter.setValue(p->recenter != 0);
  if (p->use_term_crit) {
    auto tc = std::make_shared<ICPTerminationCriteriaScore3D>();
    tc->param_d_pert.setValue(p->term_d_pert);
    tc->param_d_score.setValue(p->term_d_score);
    al->param_term_crit.setValue(tc);
  }
  al->setFixed(fixed);
  al->setMoving(moving);
  memcpy(al->X, X0, sizeof(al->X));
  memcpy(al->Xf, Xf, sizeof(al->Xf));
  for (int i = 0; i < 6; ++i) al->v[i] = v0 ? v0[i] : 0.0;
  *result = on_device ? al->computeOnDevice(stats_mode != 0) : al->compute(stats_mode != 0);
  memcpy(X_out, al->X, sizeof(al->X));
  if (v_out) memcpy(v_out, al->v, sizeof(al->v));
  *n_iter = (int)al->stats_log.size();
  auto put = [](kdh_stats* o, const ICPStats& s) {
    o->num_inliers = s.num_inliers;
    o->num_outliers = s.num_outliers;
    o->num_bad = s.num_bad;
    o->num_err_normal = s.num_err_normal;
    o->num_err_fn = s.num_err_fn;
    o->pad = 0;
    o->chi2 = s.chi2;
    o->chi2_kernelized = s.chi2_kernelized;
    o->pert_pose_norm = s.pert_pose_norm;
    o->pert_vel_norm = s.pert_vel_norm;
  };
  if (last) put(last, al->stats);
  if (log)
    for (size_t i = 0; i < al->stats_log.size() && (int)i < std::max(p->max_iterations, 1); ++i)
      put(log + i, al->stats_log[i]);
  return KDH_OK;
}

}  // namespace

extern "C" {

// Comma-separated registered type names; returns the needed buffer size when
// buf is NULL.
int kdha_registered_types(char* buf, int len) {
  std::string s;
  for (auto& n : srrg2_core::TypeRegistry::instance().names()) s += (s.empty() ? "" : ",") + n;
  if (!buf || len <= 0) return (int)s.size() + 1;
  strncpy(buf, s.c_str(), (size_t)len - 1);
  buf[len - 1] = 0;
  return 0;
}

// Creates `type_name` through the registry, configures it through its PARAM
// properties and runs compute() (on_device = 0: the plain per-call path) or
// computeOnDevice() on trees uploaded with kdh_tree_upload.  stats_log (may be
// NULL) receives stats_log, max(max_iterations, 1) entries at most.
int kdha_align(const char* type_name, kdh_ctx* ctx, const kdh_tree* fixed, const kdh_tree* moving,
               const double X0[12], const double v0[6], const double Xf[12], const kdh_params* p, int on_device,
               int stats_mode, double X_out[12], double v_out[6], int* result, int* n_iter, kdh_stats* last,
               kdh_stats* stats_log) {
  if (!type_name || !ctx || !fixed || !moving || !X0 || !Xf || !p || !X_out || !result || !n_iter)
    return KDH_ERR_INVALID;
  try {
    auto obj = srrg2_core::TypeRegistry::instance().create(type_name);
    if (auto* a = dynamic_cast<ICPHIP3D*>(obj.get()))
      return run_aligner(a, ctx, fixed, moving, X0, v0, Xf, p, on_device, stats_mode, X_out, v_out, result, n_iter,
                         last, stats_log);
    if (auto* c = dynamic_cast<CTICPHIP3D*>(obj.get()))
      return run_aligner(c, ctx, fixed, moving, X0, v0, Xf, p, on_device, stats_mode, X_out, v_out, result, n_iter,
                         last, stats_log);
    return KDH_ERR_INVALID;
  } catch (const std::exception&) {
    return KDH_ERR_INVALID;
  }
}

// Creates CTICPHIP3D through the registry and runs buildQuadraticFormDual
// at the given states; the blocks come back in a kdh_dual_form.
int kdha_build_dual(kdh_ctx* ctx, const kdh_tree* fixed, const kdh_tree* moving, const double X_moving[12],
                    const double v_moving[6], const double X_fixed[12], const double v_fixed[6], const kdh_params* p,
                    int stats_mode, kdh_dual_form* out) {
  if (!ctx || !fixed || !moving || !X_moving || !v_moving || !X_fixed || !v_fixed || !p || !out)
    return KDH_ERR_INVALID;
  try {
    auto obj = srrg2_core::TypeRegistry::instance().create("CTICPHIP3D");
    auto* c = dynamic_cast<CTICPHIP3D*>(obj.get());
    if (!c) return KDH_ERR_INVALID;
    c->setContext(ctx);
    c->param_kernel_threshold.setValue(p->kernel_threshold);
    c->param_damping.setValue(p->damping);
    c->param_max_error.setValue(p->max_error);
    c->param_min_normal_cos.setValue(p->min_normal_cos);
    c->param_min_good_measurements.setValue(p->min_good_measurements);
    c->param_max_iterations.setValue(p->max_iterations);
    c->param_velocity_damping.setValue(p->velocity_damping);
    c->param_recenter.setValue(p->recenter != 0);
    c->setFixed(fixed);
    c->setMoving(moving);
    memcpy(c->X, X_moving, sizeof(c->X));
    memcpy(c->v, v_moving, sizeof(c->v));
    memcpy(c->Xf, X_fixed, sizeof(c->Xf));
    memcpy(c->vf, v_fixed, sizeof(c->vf));
    c->buildQuadraticFormDual(stats_mode != 0);
    memcpy(out->H, c->H, sizeof(out->H));
    memcpy(out->b, c->b, sizeof(out->b));
    memcpy(out->H_dual_vel_B, c->H_dual_vel_B, sizeof(out->H_dual_vel_B));
    memcpy(out->b_dual_vel_B, c->b_dual_vel_B, sizeof(out->b_dual_vel_B));
    memcpy(out->H_dual_cross_pose_vel_B, c->H_dual_cross_pose_vel_B, sizeof(out->H_dual_cross_pose_vel_B));
    memcpy(out->H_dual_cross_vel_AB, c->H_dual_cross_vel_AB, sizeof(out->H_dual_cross_vel_AB));
    const ICPStats& s = c->stats;
    memset(&out->stats, 0, sizeof(out->stats));
    out->stats.num_inliers = s.num_inliers;
    out->stats.num_outliers = s.num_outliers;
    out->stats.num_bad = s.num_bad;
    out->stats.num_err_normal = s.num_err_normal;
    out->stats.num_err_fn = s.num_err_fn;
    out->stats.chi2 = s.chi2;
    out->stats.chi2_kernelized = s.chi2_kernelized;
    return KDH_OK;
  } catch (const std::exception&) {
    return KDH_ERR_INVALID;
  }
}

}  // extern "C"
