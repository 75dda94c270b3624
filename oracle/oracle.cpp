// ORACLE -- test infrastructure only.  Nothing in the product links or calls
// this code: only tests/, the smoke check and the CPU-baseline leg of bench.py
// load it, as the checker and as the reference CPU timing.
//
// A CPU restatement of the reference aligners ICP_CPU_ / CTICP_CPU_
// (rvp-group/kd_slam2), written without Eigen and with the reference's
// evaluation order made explicit:
//   ICP_::updateCache / oneRound / quadraticTerm
//       kd_slam2/src/lib/kd_slam/icp/icp_impl_.h:16-21, 53-90, 103-152
//   ICP_CPU_::_buildQuadraticForm (<= 8 std::async chunks, StableAdder_)
//       kd_slam2/src/lib/kd_slam/icp/icp_cpu_impl_.h:28-98
//   ICPWsItemDiag_ / ICPWsItemCross_ / ICPStats_
//       icp/icp_ws_item_.h:14-91, icp/icp_stats_.h:10-102
//   StableAdder_  utils/stable_adder_.h:21-74
//   CTICP_::oneRound / quadraticTerm, CTICP_CPU_::_buildQuadraticForm
//       cticp/cticp_impl_.h:48-92, 107-181; cticp/cticp_cpu_impl_.h:27-126
//   3D traits  d3/icp_.h:41-60, d3/cticp_impl_.h:11-39, 76-85
//   geometry   utils/geometry_.h:11-37, utils/geometry_3d_impl_.h:22-46, 143-194
//   compute()  icp/tree_aligner_base_.h:100-115, pose prior :139-152
// Compiled with -ffp-contract=off for parity (the decision expressions must
// not be fused).  The single shared piece with the product is the sine/cosine
// of the CT deskew (kdh_sincos.h): host libm and the device math library
// differ in the last bit, so both sides call one function; it is pinned
// against libm separately.
//
// Every function also has an FP64 accumulation twin: the same discrete
// decisions, with the per-leaf error and Jacobian re-evaluated in FP64 from the
// same FP32 inputs and summed in FP64, as the error yardstick for H and b.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <deque>
#include <future>
#include <limits>
#include <type_traits>
#include <vector>

#include "../kd_slam2_amd/csrc/kdh_sincos.h"

namespace orc {

// ---- Node_<Point3fTraits> memory layout (node_.h:12-43) --------------------
struct Node {
  double dts_mean;
  double dts_covariance;
  float mean[3];
  float direction[3];
  uint64_t idx_start;
  uint64_t idx_end;
  int32_t idx_left;
  int32_t idx_right;
  bool isLeaf() const { return idx_left < 0 && idx_right < 0; }
};
static_assert(sizeof(Node) == 64, "Node_ layout");

// ---- small Eigen-free types --------------------------------------------------
template <typename T> struct Vec3 { T v[3]; T& operator[](int i) { return v[i]; } T operator[](int i) const { return v[i]; } };
template <typename T> struct Mat3 { T m[3][3]; };
template <typename T> struct Iso3 { Mat3<T> R; Vec3<T> t; };

template <typename T> static Vec3<T> mk(T a, T b, T c) { Vec3<T> r; r.v[0] = a; r.v[1] = b; r.v[2] = c; return r; }
template <typename T> static Vec3<T> add(Vec3<T> a, Vec3<T> b) { return mk(a[0] + b[0], a[1] + b[1], a[2] + b[2]); }
template <typename T> static Vec3<T> sub(Vec3<T> a, Vec3<T> b) { return mk(a[0] - b[0], a[1] - b[1], a[2] - b[2]); }
// Eigen dot for 3-vectors: (a0 b0 + a1 b1) + a2 b2
template <typename T> static T dot(Vec3<T> a, Vec3<T> b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }
template <typename T> static Vec3<T> cross(Vec3<T> a, Vec3<T> b) {
  return mk(a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]);
}
template <typename T> static Vec3<T> matvec(const Mat3<T>& M, Vec3<T> p) {
  Vec3<T> r;
  for (int i = 0; i < 3; ++i) r[i] = (M.m[i][0] * p[0] + M.m[i][1] * p[1]) + M.m[i][2] * p[2];
  return r;
}
template <typename T> static Mat3<T> matmul(const Mat3<T>& A, const Mat3<T>& B) {
  Mat3<T> C;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) C.m[i][j] = (A.m[i][0] * B.m[0][j] + A.m[i][1] * B.m[1][j]) + A.m[i][2] * B.m[2][j];
  return C;
}
template <typename T> static Mat3<T> transp(const Mat3<T>& A) {
  Mat3<T> C;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) C.m[i][j] = A.m[j][i];
  return C;
}
template <typename T> static Mat3<T> eye() {
  Mat3<T> C;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) C.m[i][j] = (i == j) ? T(1) : T(0);
  return C;
}
template <typename T> static Mat3<T> skew(Vec3<T> v) {
  Mat3<T> m;
  m.m[0][0] = 0;     m.m[0][1] = -v[2]; m.m[0][2] = v[1];
  m.m[1][0] = v[2];  m.m[1][1] = 0;     m.m[1][2] = -v[0];
  m.m[2][0] = -v[1]; m.m[2][1] = v[0];  m.m[2][2] = 0;
  return m;
}
// applyIsometry_ / applyRotation_ (geometry_.h:11-37): explicit loop order.
template <typename T> static Vec3<T> applyIsometry(const Iso3<T>& X, Vec3<T> p) {
  Vec3<T> r = X.t;
  for (int i = 0; i < 3; ++i)
    for (int c = 0; c < 3; ++c) r[i] += X.R.m[i][c] * p[c];
  return r;
}
template <typename T> static Vec3<T> applyRotation(const Iso3<T>& X, Vec3<T> p) {
  Vec3<T> r = mk(T(0), T(0), T(0));
  for (int i = 0; i < 3; ++i)
    for (int c = 0; c < 3; ++c) r[i] += X.R.m[i][c] * p[c];
  return r;
}
// Eigen Isometry product and inverse.
template <typename T> static Iso3<T> compose(const Iso3<T>& A, const Iso3<T>& B) {
  Iso3<T> C;
  C.R = matmul(A.R, B.R);
  C.t = add(matvec(A.R, B.t), A.t);
  return C;
}
template <typename T> static Iso3<T> inverse(const Iso3<T>& A) {
  Iso3<T> C;
  C.R = transp(A.R);
  Vec3<T> t = matvec(C.R, A.t);
  C.t = mk(-t[0], -t[1], -t[2]);
  return C;
}
template <typename T, typename S> static Iso3<T> from34(const S* a) {
  Iso3<T> X;
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) X.R.m[r][c] = T(a[r * 4 + c]);
    X.t[r] = T(a[r * 4 + 3]);
  }
  return X;
}
template <typename T, typename S> static void to34(const Iso3<T>& X, S* a) {
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) a[r * 4 + c] = S(X.R.m[r][c]);
    a[r * 4 + 3] = S(X.t[r]);
  }
}
template <typename To, typename From> static Iso3<To> cast(const Iso3<From>& A) {
  Iso3<To> C;
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) C.R.m[r][c] = To(A.R.m[r][c]);
    C.t[r] = To(A.t[r]);
  }
  return C;
}

// ---- rodrigues / expmap / dRxp (geometry_3d_impl_.h) -------------------------
static void sincos_f(float a, float* s, float* c) { kdh::sincos_shared(a, s, c); }
static void sincos_f(double a, double* s, double* c) { *s = sin(a); *c = cos(a); }

template <typename T> static Mat3<T> rodrigues(Vec3<T> w) {
  T angle = sqrt((w[0] * w[0] + w[1] * w[1]) + w[2] * w[2]);
  if (angle < T(1e-9)) return eye<T>();
  T inv = T(1.) / angle;
  Vec3<T> n = mk(w[0] * inv, w[1] * inv, w[2] * inv);
  Mat3<T> S = skew(n);
  Mat3<T> S2 = matmul(S, S);
  T s, c;
  sincos_f(angle, &s, &c);
  Mat3<T> R;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) R.m[i][j] = (T(i == j ? 1 : 0) + s * S.m[i][j]) + (T(1.) - c) * S2.m[i][j];
  return R;
}

template <typename T> static Iso3<T> expmap(const T* v) {
  Iso3<T> X;
  X.t = mk(v[0], v[1], v[2]);
  X.R = rodrigues(mk(v[3], v[4], v[5]));
  return X;
}

template <typename T> static Vec3<T> logSO3(const Mat3<T>& R) {
  T tr = (R.m[0][0] + R.m[1][1]) + R.m[2][2];
  T cos_a = T(0.5) * (tr - T(1));
  cos_a = std::max(T(-1), std::min(T(1), cos_a));
  T angle = acos(cos_a);
  Vec3<T> ax = mk(R.m[2][1] - R.m[1][2], R.m[0][2] - R.m[2][0], R.m[1][0] - R.m[0][1]);
  if (angle < T(1e-7)) return mk(ax[0] * T(0.5), ax[1] * T(0.5), ax[2] * T(0.5));
  T f = angle / (T(2) * sin(angle));
  return mk(ax[0] * f, ax[1] * f, ax[2] * f);
}

// d(R(w) u)/dw (Gallego & Yezzi, geometry_3d_impl_.h:143-174)
template <typename T> static void dRxp(Vec3<T>& ur, Mat3<T>& dR, Vec3<T> w, Vec3<T> u) {
  T angle = sqrt((w[0] * w[0] + w[1] * w[1]) + w[2] * w[2]);
  if (angle < T(1e-9)) {
    ur = u;
    Mat3<T> s = skew(u);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) dR.m[i][j] = -s.m[i][j];
    return;
  }
  T inv = T(1.) / angle;
  Vec3<T> n = mk(w[0] * inv, w[1] * inv, w[2] * inv);
  Mat3<T> S = skew(n);
  Mat3<T> S2 = matmul(S, S);
  T s, c;
  sincos_f(angle, &s, &c);
  Mat3<T> R;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) R.m[i][j] = (T(i == j ? 1 : 0) + s * S.m[i][j]) + (T(1.) - c) * S2.m[i][j];
  ur = matvec(R, u);
  Mat3<T> Rt = transp(R);
  for (int i = 0; i < 3; ++i) Rt.m[i][i] -= T(1);
  Mat3<T> B = matmul(Rt, S);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) B.m[i][j] = n[i] * n[j] + B.m[i][j] * inv;
  Mat3<T> A = matmul(R, skew(u));
  Mat3<T> D = matmul(A, B);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) dR.m[i][j] = -D.m[i][j];
}

// ---- StableAdder_ (stable_adder_.h:21-74) ----------------------------------
template <typename T, int N> struct Packed {
  T a[N];
  void zero() { for (int i = 0; i < N; ++i) a[i] = T(0); }
  Packed& operator+=(const Packed& o) { for (int i = 0; i < N; ++i) a[i] += o.a[i]; return *this; }
};

template <typename V, int BATCH = 1024, int LEVELS = 5> struct StableAdder {
  V value[LEVELS];
  size_t count[LEVELS];
  StableAdder() { reset(); }
  void reset() {
    for (int i = 0; i < LEVELS; ++i) {
      value[i].zero();
      count[i] = 0;
    }
  }
  void add(V v) {
    int i = 0;
    while (i < LEVELS) {
      value[i] += v;
      ++count[i];
      if (count[i] < (size_t)BATCH) break;
      v = value[i];
      value[i].zero();
      count[i] = 0;
      ++i;
    }
  }
  V sum() const {
    V acc;
    acc.zero();
    for (int i = 0; i < LEVELS; ++i) acc += value[i];
    return acc;
  }
};

// ---- the aligner -------------------------------------------------------------
struct Params {
  float kernel_threshold, max_error, min_normal_cos, damping, velocity_damping;
  int min_good_measurements, max_iterations, use_term_crit;
  float d_pert, d_score;
};

// ICPStats_ counters + chi2 sums, as a packed float vector for the adder
// (counters are summed as floats: exact up to 2^24 leaves per adder level).
enum { S_IN = 0, S_OUT, S_BAD, S_NORM, S_FN, S_CHI2, S_CHI2K, S_N };

struct TermOut {
  int outcome;        // 0 inlier 1 outlier 2 bad 3 err_normal 4 err_fn
  int leaf;           // fixed leaf index or -1
  int depth;
  float stats[S_N];
  float acc[90];      // ICP: 27 used, CT: 90
  double acc64[90];   // FP64 twin
  double chi2_64, chi2k_64;
};

struct Cache {
  Iso3<float> Xm, Xf, Xmf;
  float v[6];
  // State the FP64 twin evaluates at: the FP32 cache widened (s64 false) or
  // the FP64 state the cache was rounded from (s64 true).
  bool s64 = false;
  Iso3<double> Xm64, Xf64;
  double v64[6];
};

// Descent (icp_impl_.h:116-121; whichSide node_.h:54-56)
static int descend(const Node* fixed, Vec3<float> q, int* depth) {
  int aux = 0;
  int d = 0;
  while (aux >= 0 && !fixed[aux].isLeaf()) {
    const Node& n = fixed[aux];
    Vec3<float> diff = sub(q, mk(n.mean[0], n.mean[1], n.mean[2]));
    bool left = dot(mk(n.direction[0], n.direction[1], n.direction[2]), diff) > 0;
    aux = left ? n.idx_left : n.idx_right;
    ++d;
  }
  *depth = d;
  return aux;
}

template <bool CT>
static void quadraticTerm(TermOut& o, const Node* fixed, const Node& leafnode, const Cache& C, const Params& P,
                          bool stats_mode) {
  for (int i = 0; i < S_N; ++i) o.stats[i] = 0.0f;
  for (int i = 0; i < 90; ++i) {
    o.acc[i] = 0.0f;
    o.acc64[i] = 0.0;
  }
  o.chi2_64 = o.chi2k_64 = 0.0;
  Vec3<float> mean = mk(leafnode.mean[0], leafnode.mean[1], leafnode.mean[2]);
  Vec3<float> dir = mk(leafnode.direction[0], leafnode.direction[1], leafnode.direction[2]);
  const float dt = (float)leafnode.dts_mean;
  Vec3<float> wmean = mean, wdir = dir;
  if (CT) {
    // CTICP_3D_Traits::applyVelocities (d3/cticp_impl_.h:76-85)
    Vec3<float> td = mk(C.v[0] * dt, C.v[1] * dt, C.v[2] * dt);
    Vec3<float> rdl = mk(C.v[3] * dt, C.v[4] * dt, C.v[5] * dt);
    Mat3<float> R = rodrigues(rdl);
    wmean = add(matvec(R, mean), td);
    wdir = matvec(R, dir);
  }
  const Vec3<float> p_mif = applyIsometry(C.Xmf, wmean);
  int depth = 0;
  const int aux = descend(fixed, p_mif, &depth);
  o.depth = depth;
  o.leaf = aux;
  if (aux < 0) {
    o.outcome = 2;
    o.stats[S_BAD] = 1;
    return;
  }
  const Node& fl = fixed[aux];
  const Vec3<float> pf = mk(fl.mean[0], fl.mean[1], fl.mean[2]);
  const Vec3<float> nf = mk(fl.direction[0], fl.direction[1], fl.direction[2]);
  const Vec3<float> p_m = applyIsometry(C.Xm, wmean);
  const Vec3<float> n_m = applyRotation(C.Xm, wdir);
  const Vec3<float> p_f = applyIsometry(C.Xf, pf);
  const Vec3<float> n_f = applyRotation(C.Xf, nf);
  // errorAndJacobian (d3/icp_.h:41-60)
  if (dot(n_f, n_m) < P.min_normal_cos) {
    o.outcome = 3;
    o.stats[S_NORM] = 1;
    return;
  }
  const Vec3<float> nsum = add(n_m, n_f);
  const float e = dot(sub(p_m, p_f), nsum);
  if (fabsf(e) > P.max_error) {
    o.outcome = 4;
    o.stats[S_FN] = 1;
    return;
  }
  const Vec3<float> jr = add(cross(p_m, n_f), cross(p_f, n_m));
  float J[6] = {nsum[0], nsum[1], nsum[2], jr[0], jr[1], jr[2]};
  float Jv[6] = {0, 0, 0, 0, 0, 0};
  if (CT) {
    // errorAndJacobianVelocityA (d3/cticp_impl_.h:11-39)
    float ma[6];
    for (int i = 0; i < 6; ++i) ma[i] = C.v[i] * dt;
    Vec3<float> pa_m, na_m;
    Mat3<float> dRp, dRn;
    dRxp(pa_m, dRp, mk(ma[3], ma[4], ma[5]), mean);
    dRxp(na_m, dRn, mk(ma[3], ma[4], ma[5]), dir);
    pa_m = add(pa_m, mk(ma[0], ma[1], ma[2]));
    const Mat3<float>& Ra = C.Xm.R;
    Vec3<float> pa_g = add(matvec(Ra, pa_m), C.Xm.t);
    Vec3<float> na_g = matvec(Ra, na_m);
    Vec3<float> n = add(na_g, n_f);
    Vec3<float> dp = sub(pa_g, p_f);
    Mat3<float> RdRp = matmul(Ra, dRp), RdRn = matmul(Ra, dRn);
    for (int c = 0; c < 3; ++c) {
      Jv[c] = (n[0] * Ra.m[0][c] + n[1] * Ra.m[1][c]) + n[2] * Ra.m[2][c];
      float a = (n[0] * RdRp.m[0][c] + n[1] * RdRp.m[1][c]) + n[2] * RdRp.m[2][c];
      float b = (dp[0] * RdRn.m[0][c] + dp[1] * RdRn.m[1][c]) + dp[2] * RdRn.m[2][c];
      Jv[3 + c] = a + b;
    }
    for (int i = 0; i < 6; ++i) Jv[i] *= dt;
  }
  const float chi2 = e * e;
  float gamma = 1;
  if (chi2 > P.kernel_threshold) {
    o.outcome = 1;
    o.stats[S_OUT] = 1;
    gamma = stats_mode ? 0.f : sqrtf(P.kernel_threshold / chi2);
  } else {
    o.outcome = 0;
    o.stats[S_IN] = 1;
  }
  o.stats[S_CHI2] = chi2;
  o.stats[S_CHI2K] = chi2 * gamma;
  int k = 0;
  for (int r = 0; r < 6; ++r)
    for (int c = r; c < 6; ++c, ++k) o.acc[k] = J[r] * J[c] * gamma;
  for (int r = 0; r < 6; ++r, ++k) o.acc[k] = J[r] * e * gamma;
  if (CT) {
    for (int r = 0; r < 6; ++r)
      for (int c = r; c < 6; ++c, ++k) o.acc[k] = Jv[r] * Jv[c] * gamma;
    for (int r = 0; r < 6; ++r, ++k) o.acc[k] = Jv[r] * e * gamma;
    for (int r = 0; r < 6; ++r)
      for (int c = 0; c < 6; ++c, ++k) o.acc[k] = J[r] * Jv[c] * gamma;
  }

  // FP64 twin: same inputs and decisions, FP64 arithmetic for e, J, gamma.
  {
    Iso3<double> Xm = C.s64 ? C.Xm64 : cast<double>(C.Xm), Xf = C.s64 ? C.Xf64 : cast<double>(C.Xf);
    double v64[6];
    for (int i = 0; i < 6; ++i) v64[i] = C.s64 ? C.v64[i] : (double)C.v[i];
    Vec3<double> wm = mk((double)wmean[0], (double)wmean[1], (double)wmean[2]);
    Vec3<double> wd = mk((double)wdir[0], (double)wdir[1], (double)wdir[2]);
    if (CT) {
      double dtd = (double)dt;
      double vv[6];
      for (int i = 0; i < 6; ++i) vv[i] = v64[i] * dtd;
      Mat3<double> R = rodrigues(mk(vv[3], vv[4], vv[5]));
      Vec3<double> md = mk((double)mean[0], (double)mean[1], (double)mean[2]);
      Vec3<double> dd = mk((double)dir[0], (double)dir[1], (double)dir[2]);
      wm = add(matvec(R, md), mk(vv[0], vv[1], vv[2]));
      wd = matvec(R, dd);
    }
    Vec3<double> pm = applyIsometry(Xm, wm), nm = applyRotation(Xm, wd);
    Vec3<double> pfd = applyIsometry(Xf, mk((double)pf[0], (double)pf[1], (double)pf[2]));
    Vec3<double> nfd = applyRotation(Xf, mk((double)nf[0], (double)nf[1], (double)nf[2]));
    Vec3<double> ns = add(nm, nfd);
    double ed = dot(sub(pm, pfd), ns);
    Vec3<double> jrd = add(cross(pm, nfd), cross(pfd, nm));
    double Jd[6] = {ns[0], ns[1], ns[2], jrd[0], jrd[1], jrd[2]};
    double Jvd[6] = {0, 0, 0, 0, 0, 0};
    if (CT) {
      double dtd = (double)dt;
      double ma[6];
      for (int i = 0; i < 6; ++i) ma[i] = v64[i] * dtd;
      Vec3<double> pa_m, na_m;
      Mat3<double> dRp, dRn;
      Vec3<double> md = mk((double)mean[0], (double)mean[1], (double)mean[2]);
      Vec3<double> dd = mk((double)dir[0], (double)dir[1], (double)dir[2]);
      dRxp(pa_m, dRp, mk(ma[3], ma[4], ma[5]), md);
      dRxp(na_m, dRn, mk(ma[3], ma[4], ma[5]), dd);
      pa_m = add(pa_m, mk(ma[0], ma[1], ma[2]));
      const Mat3<double>& Ra = Xm.R;
      Vec3<double> pa_g = add(matvec(Ra, pa_m), Xm.t);
      Vec3<double> na_g = matvec(Ra, na_m);
      Vec3<double> n = add(na_g, nfd);
      Vec3<double> dp = sub(pa_g, pfd);
      Mat3<double> RdRp = matmul(Ra, dRp), RdRn = matmul(Ra, dRn);
      for (int c = 0; c < 3; ++c) {
        Jvd[c] = (n[0] * Ra.m[0][c] + n[1] * Ra.m[1][c]) + n[2] * Ra.m[2][c];
        double a = (n[0] * RdRp.m[0][c] + n[1] * RdRp.m[1][c]) + n[2] * RdRp.m[2][c];
        double b = (dp[0] * RdRn.m[0][c] + dp[1] * RdRn.m[1][c]) + dp[2] * RdRn.m[2][c];
        Jvd[3 + c] = a + b;
      }
      for (int i = 0; i < 6; ++i) Jvd[i] *= dtd;
    }
    double c2 = ed * ed;
    double g = 1.0;
    if (o.outcome == 1) g = stats_mode ? 0.0 : sqrt((double)P.kernel_threshold / c2);
    o.chi2_64 = c2;
    o.chi2k_64 = c2 * g;
    int kk = 0;
    for (int r = 0; r < 6; ++r)
      for (int c = r; c < 6; ++c, ++kk) o.acc64[kk] = Jd[r] * Jd[c] * g;
    for (int r = 0; r < 6; ++r, ++kk) o.acc64[kk] = Jd[r] * ed * g;
    if (CT) {
      for (int r = 0; r < 6; ++r)
        for (int c = r; c < 6; ++c, ++kk) o.acc64[kk] = Jvd[r] * Jvd[c] * g;
      for (int r = 0; r < 6; ++r, ++kk) o.acc64[kk] = Jvd[r] * ed * g;
      for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c, ++kk) o.acc64[kk] = Jd[r] * Jvd[c] * g;
    }
  }
}

// ---- _buildQuadraticForm (icp_cpu_impl_.h:28-98, cticp_cpu_impl_.h:27-126) ---

struct BuildOut {
  float acc[90];       // summed per-leaf packs
  float stats[S_N];
  double acc64[90];
  double chi2_64, chi2k_64;
};

template <bool CT>
static BuildOut bqf_chunk(const Node* fixed, const Node* moving, const uint64_t* leaves, size_t i0, size_t i1,
                          const Cache& C, const Params& P, bool stats_mode, int32_t* corr, int32_t* depth) {
  constexpr int NA = CT ? 90 : 27;
  StableAdder<Packed<float, S_N>> sa;
  // The reference keeps one adder per workspace item (diag pose, diag vel,
  // cross); the items are independent element-wise sums, so one adder over
  // the concatenation gives identical results.
  StableAdder<Packed<float, NA>> da;
  BuildOut r;
  for (int i = 0; i < 90; ++i) r.acc64[i] = 0.0;
  r.chi2_64 = r.chi2k_64 = 0.0;
  TermOut o;
  for (size_t i = i0; i < i1; ++i) {
    quadraticTerm<CT>(o, fixed, moving[leaves[i]], C, P, stats_mode);
    Packed<float, S_N> s;
    for (int k = 0; k < S_N; ++k) s.a[k] = o.stats[k];
    Packed<float, NA> d;
    for (int k = 0; k < NA; ++k) d.a[k] = o.acc[k];
    sa.add(s);
    da.add(d);
    for (int k = 0; k < NA; ++k) r.acc64[k] += o.acc64[k];
    r.chi2_64 += o.chi2_64;
    r.chi2k_64 += o.chi2k_64;
    if (corr) corr[i] = o.leaf;
    if (depth) depth[i] = o.depth;
  }
  Packed<float, S_N> s = sa.sum();
  Packed<float, NA> d = da.sum();
  for (int k = 0; k < S_N; ++k) r.stats[k] = s.a[k];
  for (int k = 0; k < 90; ++k) r.acc[k] = k < NA ? d.a[k] : 0.0f;
  return r;
}

struct Form {
  int D;
  float H[144], b[12];
  double H64[144], b64[12];
  float stats[S_N];
  double chi2_64, chi2k_64;
};

static void unpack_diag(const float* acc, float* H, int ld, int o, float* b) {
  int k = 0;
  for (int r = 0; r < 6; ++r)
    for (int c = r; c < 6; ++c, ++k) H[(o + r) * ld + o + c] = H[(o + c) * ld + o + r] = acc[k];
  for (int r = 0; r < 6; ++r, ++k) b[o + r] = acc[k];
}
static void unpack_diag64(const double* acc, double* H, int ld, int o, double* b) {
  int k = 0;
  for (int r = 0; r < 6; ++r)
    for (int c = r; c < 6; ++c, ++k) H[(o + r) * ld + o + c] = H[(o + c) * ld + o + r] = acc[k];
  for (int r = 0; r < 6; ++r, ++k) b[o + r] = acc[k];
}

template <bool CT>
static void assemble(const BuildOut& r, Form& f) {
  constexpr int D = CT ? 12 : 6;
  f.D = D;
  for (int i = 0; i < 144; ++i) {
    f.H[i] = 0;
    f.H64[i] = 0;
  }
  unpack_diag(r.acc, f.H, D, 0, f.b);
  unpack_diag64(r.acc64, f.H64, D, 0, f.b64);
  if (CT) {
    unpack_diag(r.acc + 27, f.H, D, 6, f.b);
    unpack_diag64(r.acc64 + 27, f.H64, D, 6, f.b64);
    for (int i = 0; i < 6; ++i)
      for (int j = 0; j < 6; ++j) {
        f.H[i * D + 6 + j] = f.H[(6 + j) * D + i] = r.acc[54 + i * 6 + j];
        f.H64[i * D + 6 + j] = f.H64[(6 + j) * D + i] = r.acc64[54 + i * 6 + j];
      }
  }
  for (int k = 0; k < S_N; ++k) f.stats[k] = r.stats[k];
  f.chi2_64 = r.chi2_64;
  f.chi2k_64 = r.chi2k_64;
}

// nt = min(num_threads, 8) contiguous chunks of n/nt leaves (the first n % nt
// chunks one longer), summed in chunk order; nt <= 1 runs one chunk.
template <bool CT>
static void build_form(const Node* fixed, int64_t n_fixed, const Node* moving, const uint64_t* leaves,
                       int64_t n_leaves, const Cache& C, const Params& P, bool stats_mode, int num_threads,
                       Form& f, int32_t* corr, int32_t* depth) {
  const int MAX_THREADS = 8;
  int nt = std::min(num_threads, MAX_THREADS);
  BuildOut tot;
  if (n_fixed <= 0) {
    // An empty fixed tree contributes nothing (the reference GPU kernel returns
    // before any term; the CPU path would read node 0 of an empty array).
    memset(&tot, 0, sizeof(tot));
    for (int64_t i = 0; i < n_leaves; ++i) {
      if (corr) corr[i] = -1;
      if (depth) depth[i] = 0;
    }
  } else if (nt <= 1) {
    tot = bqf_chunk<CT>(fixed, moving, leaves, 0, (size_t)n_leaves, C, P, stats_mode, corr, depth);
  } else {
    std::future<BuildOut> thr[MAX_THREADS];
    size_t n = (size_t)n_leaves, chunk = n / nt, rem = n % nt, start = 0;
    for (int i = 0; i < nt; ++i) {
      size_t end = start + chunk + (i < (int)rem ? 1 : 0);
      thr[i] = std::async(std::launch::async, [=] {
        return bqf_chunk<CT>(fixed, moving, leaves, start, end, C, P, stats_mode, corr, depth);
      });
      start = end;
    }
    memset(&tot, 0, sizeof(tot));
    for (int i = 0; i < nt; ++i) {
      BuildOut r = thr[i].get();
      // H += ret.H ... in chunk order, on the unpacked blocks; element-wise
      // equal to adding the packed accumulators.
      for (int k = 0; k < 90; ++k) {
        tot.acc[k] += r.acc[k];
        tot.acc64[k] += r.acc64[k];
      }
      for (int k = 0; k < S_N; ++k) tot.stats[k] += r.stats[k];
      tot.chi2_64 += r.chi2_64;
      tot.chi2k_64 += r.chi2k_64;
    }
  }
  assemble<CT>(tot, f);
}

// ---- Eigen-style LDLT with diagonal pivoting, FP32 (icp_impl_.h:82) --------
template <typename T>
static void ldlt_solve(int n, const T* A_in, const T* rhs, T* x) {
  T A[144];
  int perm[12];
  for (int i = 0; i < n * n; ++i) A[i] = A_in[i];
  for (int i = 0; i < n; ++i) perm[i] = i;
  for (int k = 0; k < n; ++k) {
    // pivot: largest remaining diagonal magnitude
    int p = k;
    T best = fabs(A[k * n + k]);
    for (int i = k + 1; i < n; ++i)
      if (fabs(A[i * n + i]) > best) {
        best = fabs(A[i * n + i]);
        p = i;
      }
    if (p != k) {
      for (int j = 0; j < n; ++j) std::swap(A[k * n + j], A[p * n + j]);
      for (int j = 0; j < n; ++j) std::swap(A[j * n + k], A[j * n + p]);
      std::swap(perm[k], perm[p]);
    }
    // A(k,k) -= sum_{j<k} L(k,j)^2 D(j) ; L(i,k) = (A(i,k) - sum L(i,j) L(k,j) D(j)) / D(k)
    T d = A[k * n + k];
    for (int j = 0; j < k; ++j) d -= A[k * n + j] * A[k * n + j] * A[j * n + j];
    A[k * n + k] = d;
    for (int i = k + 1; i < n; ++i) {
      T s = A[i * n + k];
      for (int j = 0; j < k; ++j) s -= A[i * n + j] * A[k * n + j] * A[j * n + j];
      A[i * n + k] = d != T(0) ? s / d : T(0);
    }
  }
  T y[12];
  for (int i = 0; i < n; ++i) y[i] = rhs[perm[i]];
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < i; ++j) y[i] -= A[i * n + j] * y[j];
  for (int i = 0; i < n; ++i) y[i] = A[i * n + i] != T(0) ? y[i] / A[i * n + i] : T(0);
  for (int i = n - 1; i >= 0; --i)
    for (int j = i + 1; j < n; ++j) y[i] -= A[j * n + i] * y[j];
  for (int i = 0; i < n; ++i) x[perm[i]] = y[i];
}

// ---- pose prior in FP32 (tree_aligner_base_.h:139-152, geometry_3d_impl_.h) ---
static void dlogSO3_f(const Mat3<float>& R, float J[27]) {
  for (int i = 0; i < 27; ++i) J[i] = 0;
  Vec3<float> ax = mk(R.m[2][1] - R.m[1][2], R.m[0][2] - R.m[2][0], R.m[1][0] - R.m[0][1]);
  float cos_a = 0.5f * (((R.m[0][0] + R.m[1][1]) + R.m[2][2]) - 1.0f);
  cos_a = std::max(-1.0f, std::min(1.0f, cos_a));
  float angle = acosf(cos_a);
  bool small = angle < 1e-7f;
  float f = small ? 0.5f : angle / (2.0f * sinf(angle));
  J[0 * 9 + 5] = f; J[0 * 9 + 7] = -f;
  J[1 * 9 + 6] = f; J[1 * 9 + 2] = -f;
  J[2 * 9 + 1] = f; J[2 * 9 + 3] = -f;
  if (!small) {
    float s = sinf(angle);
    float c = (s - angle * cosf(angle)) / (4.0f * s * s * s);
    for (int k = 0; k < 3; ++k) {
      J[k * 9 + 0] -= c * ax[k];
      J[k * 9 + 4] -= c * ax[k];
      J[k * 9 + 8] -= c * ax[k];
    }
  }
}

// poseToPoseError and dPoseToPoseError (geometry_3d_impl_.h:117-140):
// e = logmap(ZXi * X); J = dlogmap(P) * B with B the 12x6 derivative of
// P = ZXi * X under the left perturbation X <- exp(dx) X.
static void pose_error_f(const Iso3<float>& ZXi, const Iso3<float>& X, float e[6], float J[36]) {
  Iso3<float> P = compose(ZXi, X);
  Vec3<float> w = logSO3(P.R);
  e[0] = P.t[0]; e[1] = P.t[1]; e[2] = P.t[2];
  e[3] = w[0]; e[4] = w[1]; e[5] = w[2];
  if (!J) return;
  float B[72];
  for (int i = 0; i < 72; ++i) B[i] = 0;
  Mat3<float> Rst = matmul(ZXi.R, skew(X.t));
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      B[r * 6 + c] = ZXi.R.m[r][c];
      B[r * 6 + 3 + c] = -Rst.m[r][c];
    }
  for (int i = 0; i < 3; ++i) {
    Mat3<float> Rsc = matmul(ZXi.R, skew(mk(X.R.m[0][i], X.R.m[1][i], X.R.m[2][i])));
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) B[(3 + 3 * i + r) * 6 + 3 + c] = -Rsc.m[r][c];
  }
  float Dl[27];
  dlogSO3_f(P.R, Dl);
  for (int r = 0; r < 6; ++r)
    for (int c = 0; c < 6; ++c) {
      float s = 0;
      if (r < 3) s = B[r * 6 + c];
      else
        for (int k = 0; k < 9; ++k) s += Dl[(r - 3) * 9 + k] * B[(3 + k) * 6 + c];
      J[r * 6 + c] = s;
    }
}

static void prior_form(const Iso3<float>& ZXi, const Iso3<float>& X, const float* info, float* H6, float* b6) {
  float e[6], J[36];
  pose_error_f(ZXi, X, e, J);
  for (int r = 0; r < 6; ++r) {
    float sb = 0;
    for (int k = 0; k < 6; ++k) {
      float oe = 0;
      for (int m = 0; m < 6; ++m) oe += info[k * 6 + m] * e[m];
      sb += J[k * 6 + r] * oe;
    }
    b6[r] = sb;
    for (int c = 0; c < 6; ++c) {
      float s = 0;
      for (int k = 0; k < 6; ++k) {
        float oj = 0;
        for (int m = 0; m < 6; ++m) oj += info[k * 6 + m] * J[m * 6 + c];
        s += J[k * 6 + r] * oj;
      }
      H6[r * 6 + c] = s;
    }
  }
}

}  // namespace orc

// ---- C exports (ctypes) -------------------------------------------------------
using namespace orc;

extern "C" {

struct orc_params {
  float kernel_threshold, damping, max_error, min_normal_cos;
  int32_t min_good_measurements, max_iterations;
  float velocity_damping;
  int32_t use_term_crit;
  float term_d_pert, term_d_score;
  int32_t num_threads;
};

struct orc_stats {
  int32_t num_inliers, num_outliers, num_bad, num_err_normal, num_err_fn, pad;
  double chi2, chi2_kernelized, pert_pose_norm, pert_vel_norm;
};

static Params to_params(const orc_params* p) {
  Params P;
  P.kernel_threshold = p->kernel_threshold;
  P.max_error = p->max_error;
  P.min_normal_cos = p->min_normal_cos;
  P.damping = p->damping;
  P.velocity_damping = p->velocity_damping;
  P.min_good_measurements = p->min_good_measurements;
  P.max_iterations = p->max_iterations;
  P.use_term_crit = p->use_term_crit;
  P.d_pert = p->term_d_pert;
  P.d_score = p->term_d_score;
  return P;
}

static void stats_out(const Form& f, orc_stats* s) {
  s->num_inliers = (int)f.stats[S_IN];
  s->num_outliers = (int)f.stats[S_OUT];
  s->num_bad = (int)f.stats[S_BAD];
  s->num_err_normal = (int)f.stats[S_NORM];
  s->num_err_fn = (int)f.stats[S_FN];
  s->pad = 0;
  s->chi2 = f.stats[S_CHI2];
  s->chi2_kernelized = f.stats[S_CHI2K];
  s->pert_pose_norm = 0;
  s->pert_vel_norm = 0;
}

// One quadratic form at a given FP32 cache.  H/b: FP32 oracle (D x D, D);
// H64/b64/chi2_64: the FP64 twin; corr/depth: per moving leaf (may be NULL).
int orc_build(int ct, const void* fixed_nodes, int64_t n_fixed, const void* moving_nodes, const uint64_t* leaves,
              int64_t n_leaves, const float* Xm, const float* Xf, const float* Xmf, const float* v,
              const orc_params* p, int stats_mode, float* H, float* b, double* H64, double* b64, double* chi2_64,
              orc_stats* st, int32_t* corr, int32_t* depth) {
  Cache C;
  C.Xm = from34<float>(Xm);
  C.Xf = from34<float>(Xf);
  C.Xmf = from34<float>(Xmf);
  for (int i = 0; i < 6; ++i) C.v[i] = (ct && v) ? v[i] : 0.0f;
  Params P = to_params(p);
  Form f;
  const Node* fx = (const Node*)fixed_nodes;
  const Node* mv = (const Node*)moving_nodes;
  if (ct)
    build_form<true>(fx, n_fixed, mv, leaves, n_leaves, C, P, stats_mode != 0, p->num_threads, f, corr, depth);
  else
    build_form<false>(fx, n_fixed, mv, leaves, n_leaves, C, P, stats_mode != 0, p->num_threads, f, corr, depth);
  const int D = f.D;
  if (H) memcpy(H, f.H, sizeof(float) * D * D);
  if (b) memcpy(b, f.b, sizeof(float) * D);
  if (H64) memcpy(H64, f.H64, sizeof(double) * D * D);
  if (b64) memcpy(b64, f.b64, sizeof(double) * D);
  if (chi2_64) {
    chi2_64[0] = f.chi2_64;
    chi2_64[1] = f.chi2k_64;
  }
  if (st) stats_out(f, st);
  return 0;
}

// FP32 cache rounded once from an FP64 state (updateCache, icp_impl_.h:16-21,
// formed in FP64), remembering the state for the FP64 twin.
static Cache cache_from_state(const Iso3<double>& X, const Iso3<double>& Xf, const double* v, bool ct) {
  Cache C;
  C.Xm = cast<float>(X);
  C.Xf = cast<float>(Xf);
  C.Xmf = cast<float>(compose(inverse(Xf), X));
  C.s64 = true;
  C.Xm64 = X;
  C.Xf64 = Xf;
  for (int i = 0; i < 6; ++i) {
    C.v64[i] = (ct && v) ? v[i] : 0.0;
    C.v[i] = (float)C.v64[i];
  }
  return C;
}

// orc_build at an FP64 state: FP32 form at the rounded cache, FP64 twin at
// the state itself (X, Xf 3x4 row-major, v 6 or NULL).
int orc_build_at(int ct, const void* fixed_nodes, int64_t n_fixed, const void* moving_nodes, const uint64_t* leaves,
                 int64_t n_leaves, const double* X, const double* Xf, const double* v, const orc_params* p,
                 float* H, float* b, double* H64, double* b64, int32_t* corr) {
  Cache C = cache_from_state(from34<double>(X), from34<double>(Xf), v, ct != 0);
  Params P = to_params(p);
  Form f;
  const Node* fx = (const Node*)fixed_nodes;
  const Node* mv = (const Node*)moving_nodes;
  if (ct)
    build_form<true>(fx, n_fixed, mv, leaves, n_leaves, C, P, false, p->num_threads, f, corr, nullptr);
  else
    build_form<false>(fx, n_fixed, mv, leaves, n_leaves, C, P, false, p->num_threads, f, corr, nullptr);
  const int D = f.D;
  if (H) memcpy(H, f.H, sizeof(float) * D * D);
  if (b) memcpy(b, f.b, sizeof(float) * D);
  if (H64) memcpy(H64, f.H64, sizeof(double) * D * D);
  if (b64) memcpy(b64, f.b64, sizeof(double) * D);
  return 0;
}

// updateCache (icp_impl_.h:16-21) in FP32.
void orc_update_cache(const float* Xm, const float* Xf, float* Xmf) {
  Iso3<float> m = from34<float>(Xm), f = from34<float>(Xf);
  to34(compose(inverse(f), m), Xmf);
}

// TreeAlignerBase_::compute with ICP_/CTICP_::oneRound, FP32 throughout
// (tree_aligner_base_.h:100-115, icp_impl_.h:53-90, cticp_impl_.h:48-92).
// X: moving pose in/out (3x4 float), v: velocities in/out (CT), Xf: fixed
// pose.  prior_* may be NULL.  stats_log: max_iterations entries.
// TreeAlignerBase_::compute with ICP_/CTICP_::oneRound over n_entries fixed
// trees (tree_aligner_base_.h:100-115, icp_impl_.h:53-90, cticp_impl_.h:48-92):
// with one entry the plain buildQuadraticForm; with a fixed forest every
// entry's form is built at its own fixed pose (buildQuadraticForm(FixedEntry&),
// icp_impl_.h:30-40) and H, b and the stats are summed in entry order in FP32.
// entry_stats (may be NULL) receives each entry's stats of the last round.
static int compute_impl(int ct, int n_entries, const void* const* fixed_nodes, const int64_t* n_fixed,
                        const float* Xf, bool forest, const void* moving_nodes, const uint64_t* leaves,
                        int64_t n_leaves, float* X, float* v, const orc_params* p, const float* prior_ZXi,
                        const float* prior_info, int stats_mode, orc_stats* stats_log, int* n_log, int* result,
                        orc_stats* entry_stats) {
  Params P = to_params(p);
  Iso3<float> Xs = from34<float>(X);
  float vel[6];
  for (int i = 0; i < 6; ++i) vel[i] = (ct && v) ? v[i] : 0.0f;
  const int D = ct ? 12 : 6;
  const int max_iter = stats_mode ? 1 : P.max_iterations;
  float prev_score = 0;
  int ret = P.max_iterations;
  int nl = 0;
  const Node* mv = (const Node*)moving_nodes;
  for (int it = 0; it < max_iter; ++it) {
    float H[144] = {0}, bv[12] = {0};
    orc_stats st;
    memset(&st, 0, sizeof(st));
    for (int e = 0; e < n_entries; ++e) {
      const Iso3<float> Xfs = from34<float>(Xf + 12 * e);
      Cache C;
      C.Xm = Xs;
      C.Xf = Xfs;
      C.Xmf = compose(inverse(Xfs), Xs);
      for (int i = 0; i < 6; ++i) C.v[i] = vel[i];
      Form f;
      const Node* fx = (const Node*)fixed_nodes[e];
      if (ct)
        build_form<true>(fx, n_fixed[e], mv, leaves, n_leaves, C, P, stats_mode != 0, p->num_threads, f, nullptr,
                         nullptr);
      else
        build_form<false>(fx, n_fixed[e], mv, leaves, n_leaves, C, P, stats_mode != 0, p->num_threads, f, nullptr,
                          nullptr);
      orc_stats se;
      stats_out(f, &se);
      if (entry_stats) entry_stats[e] = se;
      if (!forest) {
        memcpy(H, f.H, sizeof(float) * D * D);
        memcpy(bv, f.b, sizeof(float) * D);
        st = se;
      } else {
        // temp_H += entry.H; temp_b += entry.b; temp_stats += entry.stats
        for (int i = 0; i < D * D; ++i) H[i] += f.H[i];
        for (int i = 0; i < D; ++i) bv[i] += f.b[i];
        st.num_inliers += se.num_inliers;
        st.num_outliers += se.num_outliers;
        st.num_bad += se.num_bad;
        st.num_err_normal += se.num_err_normal;
        st.num_err_fn += se.num_err_fn;
        st.chi2 = (float)st.chi2 + (float)se.chi2;
        st.chi2_kernelized = (float)st.chi2_kernelized + (float)se.chi2_kernelized;
      }
    }
    if (prior_ZXi && prior_info) {
      float Hp[36], bp[6];
      prior_form(from34<float>(prior_ZXi), Xs, prior_info, Hp, bp);
      for (int r = 0; r < 6; ++r) {
        bv[r] += bp[r];
        for (int c = 0; c < 6; ++c) H[r * D + c] += Hp[r * 6 + c];
      }
    }
    const int good = st.num_outliers + st.num_inliers;
    if (good < P.min_good_measurements) {
      stats_log[nl++] = st;
      ret = -it;
      break;
    }
    float A[144], rhs[12], dx[12];
    for (int i = 0; i < D * D; ++i) A[i] = H[i];
    for (int i = 0; i < D; ++i) {
      A[i * D + i] += (i < 6) ? P.damping : P.velocity_damping;
      rhs[i] = -bv[i];
    }
    ldlt_solve<float>(D, A, rhs, dx);
    float pn = 0, vn = 0;
    for (int i = 0; i < 6; ++i) pn += dx[i] * dx[i];
    for (int i = 6; i < D; ++i) vn += dx[i] * dx[i];
    st.pert_pose_norm = sqrtf(pn);
    st.pert_vel_norm = ct ? sqrtf(vn) : 0.0;
    stats_log[nl++] = st;
    if (!stats_mode) {
      Xs = compose(expmap<float>(dx), Xs);
      if (ct)
        for (int i = 0; i < 6; ++i) vel[i] += dx[6 + i];
    }
    if (P.use_term_crit) {
      const int nleaves = st.num_inliers + st.num_outliers + st.num_bad + st.num_err_normal + st.num_err_fn;
      const float ratio = nleaves ? (float)st.num_inliers / (float)nleaves : 0.0f;
      const float score = ratio / ((float)st.chi2_kernelized / st.num_inliers + 1e-6f);
      bool stop = false;
      if (it > 0) {
        bool pert_ok = (float)st.pert_pose_norm < P.d_pert;
        float ds = score - prev_score;
        bool score_ok = (ds > 0) && ds / (prev_score + 1e-6) < P.d_score;
        stop = pert_ok && score_ok;
      }
      prev_score = score;
      if (stop) {
        ret = it;
        break;
      }
    }
  }
  to34(Xs, X);
  if (ct && v)
    for (int i = 0; i < 6; ++i) v[i] = vel[i];
  if (n_log) *n_log = nl;
  if (result) *result = ret;
  return 0;
}

int orc_compute(int ct, const void* fixed_nodes, int64_t n_fixed, const void* moving_nodes, const uint64_t* leaves,
                int64_t n_leaves, float* X, float* v, const float* Xf, const orc_params* p,
                const float* prior_ZXi, const float* prior_info, int stats_mode, orc_stats* stats_log, int* n_log,
                int* result) {
  const void* fx[1] = {fixed_nodes};
  return compute_impl(ct, 1, fx, &n_fixed, Xf, false, moving_nodes, leaves, n_leaves, X, v, p, prior_ZXi,
                      prior_info, stats_mode, stats_log, n_log, result, nullptr);
}

// compute() with a fixed forest (_fixed_forest, makeFixedEntry; icp_.h:90-95):
// n_entries fixed trees with their poses Xf (n_entries x 12).
int orc_compute_forest(int ct, int n_entries, const void* const* fixed_nodes, const int64_t* n_fixed,
                       const float* Xf, const void* moving_nodes, const uint64_t* leaves, int64_t n_leaves,
                       float* X, float* v, const orc_params* p, const float* prior_ZXi, const float* prior_info,
                       int stats_mode, orc_stats* stats_log, int* n_log, int* result, orc_stats* entry_stats) {
  if (n_entries < 1) return -1;
  return compute_impl(ct, n_entries, fixed_nodes, n_fixed, Xf, true, moving_nodes, leaves, n_leaves, X, v, p,
                      prior_ZXi, prior_info, stats_mode, stats_log, n_log, result, entry_stats);
}

// FP64 yardstick for whole alignments: the reference's FP32 discrete
// decisions (descent, gates, inlier test) at an FP32 cache rounded once from
// an FP64 state, with every accepted term re-evaluated in FP64 at that state,
// summed in FP64, solved in FP64 and the pose updated in FP64.  Rigid or CT,
// no prior.
int orc_compute64(int ct, const void* fixed_nodes, int64_t n_fixed, const void* moving_nodes, const uint64_t* leaves,
                  int64_t n_leaves, double* X, double* v, const double* Xf, const orc_params* p, int* n_iter) {
  Params P = to_params(p);
  Iso3<double> Xs = from34<double>(X);
  const Iso3<double> Xfs = from34<double>(Xf);
  double vel[6];
  for (int i = 0; i < 6; ++i) vel[i] = (ct && v) ? v[i] : 0.0;
  const int D = ct ? 12 : 6;
  int it = 0;
  for (; it < P.max_iterations; ++it) {
    Cache C = cache_from_state(Xs, Xfs, vel, ct != 0);
    Form f;
    const Node* fx = (const Node*)fixed_nodes;
    const Node* mv = (const Node*)moving_nodes;
    if (ct)
      build_form<true>(fx, n_fixed, mv, leaves, n_leaves, C, P, false, p->num_threads, f, nullptr, nullptr);
    else
      build_form<false>(fx, n_fixed, mv, leaves, n_leaves, C, P, false, p->num_threads, f, nullptr, nullptr);
    if ((int)(f.stats[S_IN] + f.stats[S_OUT]) < P.min_good_measurements) break;
    double A[144], rhs[12], dx[12];
    for (int i = 0; i < D * D; ++i) A[i] = f.H64[i];
    for (int i = 0; i < D; ++i) {
      A[i * D + i] += (i < 6) ? (double)P.damping : (double)P.velocity_damping;
      rhs[i] = -f.b64[i];
    }
    ldlt_solve<double>(D, A, rhs, dx);
    Xs = compose(expmap<double>(dx), Xs);
    if (ct)
      for (int i = 0; i < 6; ++i) vel[i] += dx[6 + i];
  }
  to34(Xs, X);
  if (ct && v)
    for (int i = 0; i < 6; ++i) v[i] = vel[i];
  if (n_iter) *n_iter = it;
  return 0;
}

}  // extern "C"

// ---- dual form (cticp_impl_.h:185-293; CPU assembly cticp_cpu_impl_.h:129-205) ----

// errorAndJacobianVelocityA (d3/cticp_impl_.h:11-39).
template <typename T>
static void vel_jacobian_A(T* J, const Iso3<T>& Ta, Vec3<T> pa_local, Vec3<T> na_local, Vec3<T> pb_global,
                           Vec3<T> nb_global, const T* va, T dt) {
  T ma[6];
  for (int i = 0; i < 6; ++i) ma[i] = va[i] * dt;
  Vec3<T> pa_moved, na_moved;
  Mat3<T> dRp, dRn;
  dRxp(pa_moved, dRp, mk(ma[3], ma[4], ma[5]), pa_local);
  dRxp(na_moved, dRn, mk(ma[3], ma[4], ma[5]), na_local);
  pa_moved = add(pa_moved, mk(ma[0], ma[1], ma[2]));
  const Mat3<T>& Ra = Ta.R;
  const Vec3<T> pa_g = add(matvec(Ra, pa_moved), Ta.t);
  const Vec3<T> na_g = matvec(Ra, na_moved);
  const Vec3<T> n = add(na_g, nb_global);
  const Vec3<T> dp = sub(pa_g, pb_global);
  const Mat3<T> RdRp = matmul(Ra, dRp), RdRn = matmul(Ra, dRn);
  for (int c = 0; c < 3; ++c) {
    J[c] = (n[0] * Ra.m[0][c] + n[1] * Ra.m[1][c]) + n[2] * Ra.m[2][c];
    const T a = (n[0] * RdRp.m[0][c] + n[1] * RdRp.m[1][c]) + n[2] * RdRp.m[2][c];
    const T b = (dp[0] * RdRn.m[0][c] + dp[1] * RdRn.m[1][c]) + dp[2] * RdRn.m[2][c];
    J[3 + c] = a + b;
  }
  for (int i = 0; i < 6; ++i) J[i] *= dt;
}

// errorAndJacobianVelocityB (d3/cticp_impl_.h:43-73): the Jacobian of the
// same error with respect to the fixed side's velocities, at its unwarped leaf.
template <typename T>
static void vel_jacobian_B(T* J, const Iso3<T>& Tb, Vec3<T> pa_global, Vec3<T> na_global, Vec3<T> pb_local,
                           Vec3<T> nb_local, const T* vb, T dt) {
  T mb[6];
  for (int i = 0; i < 6; ++i) mb[i] = vb[i] * dt;
  Vec3<T> pb_moved, nb_moved;
  Mat3<T> dRp, dRn;
  dRxp(pb_moved, dRp, mk(mb[3], mb[4], mb[5]), pb_local);
  dRxp(nb_moved, dRn, mk(mb[3], mb[4], mb[5]), nb_local);
  pb_moved = add(pb_moved, mk(mb[0], mb[1], mb[2]));
  const Mat3<T>& Rb = Tb.R;
  const Vec3<T> pb_g = add(matvec(Rb, pb_moved), Tb.t);
  const Vec3<T> nb_g = matvec(Rb, nb_moved);
  const Vec3<T> n = add(na_global, nb_g);
  const Vec3<T> dp = sub(pa_global, pb_g);
  const Mat3<T> RdRp = matmul(Rb, dRp), RdRn = matmul(Rb, dRn);
  for (int c = 0; c < 3; ++c) {
    J[c] = -((n[0] * Rb.m[0][c] + n[1] * Rb.m[1][c]) + n[2] * Rb.m[2][c]);
    const T a = (n[0] * RdRp.m[0][c] + n[1] * RdRp.m[1][c]) + n[2] * RdRp.m[2][c];
    const T b = (dp[0] * RdRn.m[0][c] + dp[1] * RdRn.m[1][c]) + dp[2] * RdRn.m[2][c];
    J[3 + c] = -a + b;
  }
  for (int i = 0; i < 6; ++i) J[i] *= dt;
}

// Per leaf: the 97 standard CT values (stats, then ICPWsItemDiag_ pose, vel,
// ICPWsItemCross_) and the dual extras: Diag(vel_B) 27, Cross(pose, vel_B) 36,
// Cross(vel_A, vel_B) 36 (icp_ws_item_.h).  kDualN = 90 + 99 packed values.
constexpr int kDualN = 189;

template <typename T>
static void pack_dual(T* out, const T* J, const T* Jv, const T* Jb, T e, T g) {
  int k = 0;
  for (int r = 0; r < 6; ++r)
    for (int c = r; c < 6; ++c, ++k) out[k] = J[r] * J[c] * g;
  for (int r = 0; r < 6; ++r, ++k) out[k] = J[r] * e * g;
  for (int r = 0; r < 6; ++r)
    for (int c = r; c < 6; ++c, ++k) out[k] = Jv[r] * Jv[c] * g;
  for (int r = 0; r < 6; ++r, ++k) out[k] = Jv[r] * e * g;
  for (int r = 0; r < 6; ++r)
    for (int c = 0; c < 6; ++c, ++k) out[k] = J[r] * Jv[c] * g;
  for (int r = 0; r < 6; ++r)
    for (int c = r; c < 6; ++c, ++k) out[k] = Jb[r] * Jb[c] * g;
  for (int r = 0; r < 6; ++r, ++k) out[k] = Jb[r] * e * g;
  for (int r = 0; r < 6; ++r)
    for (int c = 0; c < 6; ++c, ++k) out[k] = J[r] * Jb[c] * g;
  for (int r = 0; r < 6; ++r)
    for (int c = 0; c < 6; ++c, ++k) out[k] = Jv[r] * Jb[c] * g;
}

struct DualCache {
  Iso3<float> Xm, Xf, Xmf;
  float va[6], vb[6];
  Iso3<double> Xm64, Xf64;
  double va64[6], vb64[6];
};

// quadraticTermDual: the discrete part and the A side as quadraticTerm, the
// dual weight rule (inliers zeroed in stats mode, outliers always weighted,
// cticp_impl_.h:240-248) and J_vel_B at the fixed leaf's own _dts_mean.
static void quadraticTermDual(float* stats, float* acc, double* acc64, int* leaf_out, const Node* fixed,
                              const Node& leafnode, const DualCache& C, const Params& P, bool stats_mode) {
  for (int i = 0; i < S_N; ++i) stats[i] = 0.0f;
  for (int i = 0; i < kDualN; ++i) {
    acc[i] = 0.0f;
    acc64[i] = 0.0;
  }
  const Vec3<float> mean = mk(leafnode.mean[0], leafnode.mean[1], leafnode.mean[2]);
  const Vec3<float> dir = mk(leafnode.direction[0], leafnode.direction[1], leafnode.direction[2]);
  const float dt = (float)leafnode.dts_mean;
  const Vec3<float> td = mk(C.va[0] * dt, C.va[1] * dt, C.va[2] * dt);
  const Mat3<float> Rw = rodrigues(mk(C.va[3] * dt, C.va[4] * dt, C.va[5] * dt));
  const Vec3<float> wmean = add(matvec(Rw, mean), td), wdir = matvec(Rw, dir);
  const Vec3<float> p_mif = applyIsometry(C.Xmf, wmean);
  int depth = 0;
  const int aux = descend(fixed, p_mif, &depth);
  *leaf_out = aux;
  if (aux < 0) {
    stats[S_BAD] = 1;
    return;
  }
  const Node& fl = fixed[aux];
  const Vec3<float> pf = mk(fl.mean[0], fl.mean[1], fl.mean[2]);
  const Vec3<float> nf = mk(fl.direction[0], fl.direction[1], fl.direction[2]);
  const Vec3<float> p_m = applyIsometry(C.Xm, wmean), n_m = applyRotation(C.Xm, wdir);
  const Vec3<float> p_f = applyIsometry(C.Xf, pf), n_f = applyRotation(C.Xf, nf);
  if (dot(n_f, n_m) < P.min_normal_cos) {
    stats[S_NORM] = 1;
    return;
  }
  const Vec3<float> ns = add(n_m, n_f);
  const float e = dot(sub(p_m, p_f), ns);
  if (fabsf(e) > P.max_error) {
    stats[S_FN] = 1;
    return;
  }
  const Vec3<float> jr = add(cross(p_m, n_f), cross(p_f, n_m));
  const float J[6] = {ns[0], ns[1], ns[2], jr[0], jr[1], jr[2]};
  float Jv[6], Jb[6];
  vel_jacobian_A<float>(Jv, C.Xm, mean, dir, p_f, n_f, C.va, dt);
  const float dtb = (float)fl.dts_mean;
  vel_jacobian_B<float>(Jb, C.Xf, p_m, n_m, pf, nf, C.vb, dtb);
  const float chi2 = e * e;
  float gamma = stats_mode ? 0.0f : 1.0f;
  if (chi2 > P.kernel_threshold) {
    stats[S_OUT] = 1;
    gamma = sqrtf(P.kernel_threshold / chi2);
  } else {
    stats[S_IN] = 1;
  }
  stats[S_CHI2] = chi2;
  stats[S_CHI2K] = chi2 * gamma;
  pack_dual<float>(acc, J, Jv, Jb, e, gamma);

  // FP64 twin at the FP64 state, same decisions
  const Vec3<double> md = mk((double)mean[0], (double)mean[1], (double)mean[2]);
  const Vec3<double> dd = mk((double)dir[0], (double)dir[1], (double)dir[2]);
  const double dtd = (double)dt;
  const Mat3<double> Rw64 = rodrigues(mk(C.va64[3] * dtd, C.va64[4] * dtd, C.va64[5] * dtd));
  const Vec3<double> wm = add(matvec(Rw64, md), mk(C.va64[0] * dtd, C.va64[1] * dtd, C.va64[2] * dtd));
  const Vec3<double> wd = matvec(Rw64, dd);
  const Vec3<double> pfd = mk((double)pf[0], (double)pf[1], (double)pf[2]);
  const Vec3<double> nfd = mk((double)nf[0], (double)nf[1], (double)nf[2]);
  const Vec3<double> pm = applyIsometry(C.Xm64, wm), nm = applyRotation(C.Xm64, wd);
  const Vec3<double> pg = applyIsometry(C.Xf64, pfd), ng = applyRotation(C.Xf64, nfd);
  const Vec3<double> ns64 = add(nm, ng);
  const double ed = dot(sub(pm, pg), ns64);
  const Vec3<double> jrd = add(cross(pm, ng), cross(pg, nm));
  const double Jd[6] = {ns64[0], ns64[1], ns64[2], jrd[0], jrd[1], jrd[2]};
  double Jvd[6], Jbd[6];
  vel_jacobian_A<double>(Jvd, C.Xm64, md, dd, pg, ng, C.va64, dtd);
  vel_jacobian_B<double>(Jbd, C.Xf64, pm, nm, pfd, nfd, C.vb64, (double)dtb);
  double g64 = stats_mode ? 0.0 : 1.0;
  if (stats[S_OUT] != 0.0f) g64 = sqrt((double)P.kernel_threshold / (ed * ed));
  pack_dual<double>(acc64, Jd, Jvd, Jbd, ed, g64);
}

extern "C" {

// _buildQuadraticFormDual of CTICP_CPU_ (cticp_cpu_impl_.h:129-205): one
// sequential StableAdder pass over the moving leaves (the items are
// element-wise independent sums: one adder over their concatenation).
// Outputs (FP32 reference, then the FP64 twin): H 12x12, b 12, H_vel_B 6x6,
// b_vel_B 6, H_cross_pose_vel_B 6x6, H_cross_vel_AB 6x6; stats; corr.
int orc_build_dual(const void* fixed_nodes, int64_t n_fixed, const void* moving_nodes, const uint64_t* leaves,
                   int64_t n_leaves, const double* Xm, const double* vm, const double* Xf, const double* vf,
                   const orc_params* p, int stats_mode, float* out32, double* out64, orc_stats* st,
                   int32_t* corr) {
  DualCache C;
  C.Xm64 = from34<double>(Xm);
  C.Xf64 = from34<double>(Xf);
  C.Xm = cast<float>(C.Xm64);
  C.Xf = cast<float>(C.Xf64);
  C.Xmf = cast<float>(compose(inverse(C.Xf64), C.Xm64));
  for (int i = 0; i < 6; ++i) {
    C.va64[i] = vm[i];
    C.vb64[i] = vf[i];
    C.va[i] = (float)vm[i];
    C.vb[i] = (float)vf[i];
  }
  const Params P = to_params(p);
  const Node* fx = (const Node*)fixed_nodes;
  const Node* mv = (const Node*)moving_nodes;
  StableAdder<Packed<float, S_N>> sa;
  StableAdder<Packed<float, kDualN>> da;
  static thread_local double a64[kDualN];
  for (int i = 0; i < kDualN; ++i) a64[i] = 0.0;
  float stats[S_N], acc[kDualN];
  double acc64[kDualN];
  for (int64_t i = 0; i < n_leaves && n_fixed > 0; ++i) {
    int leaf = -1;
    quadraticTermDual(stats, acc, acc64, &leaf, fx, mv[leaves[i]], C, P, stats_mode != 0);
    Packed<float, S_N> s;
    for (int k = 0; k < S_N; ++k) s.a[k] = stats[k];
    Packed<float, kDualN> d;
    for (int k = 0; k < kDualN; ++k) d.a[k] = acc[k];
    sa.add(s);
    da.add(d);
    for (int k = 0; k < kDualN; ++k) a64[k] += acc64[k];
    if (corr) corr[i] = leaf;
  }
  Packed<float, S_N> s = sa.sum();
  Packed<float, kDualN> d = da.sum();
  // layout of out32 / out64: H[144], b[12], Hvb[36], bvb[6], Hcpb[36], Hcab[36] = 270
  auto unpack = [](auto* o, const auto* a) {
    using T = std::remove_reference_t<decltype(o[0])>;
    T* H = o;
    T* b = o + 144;
    T* Hvb = o + 156;
    T* bvb = o + 192;
    T* Hcpb = o + 198;
    T* Hcab = o + 234;
    for (int i = 0; i < 270; ++i) o[i] = T(0);
    int k = 0;
    for (int blk = 0; blk < 2; ++blk) {
      const int off = 6 * blk;
      for (int r = 0; r < 6; ++r)
        for (int c = r; c < 6; ++c, ++k) H[(off + r) * 12 + off + c] = H[(off + c) * 12 + off + r] = (T)a[k];
      for (int r = 0; r < 6; ++r, ++k) b[off + r] = (T)a[k];
    }
    for (int r = 0; r < 6; ++r)
      for (int c = 0; c < 6; ++c, ++k) H[r * 12 + 6 + c] = H[(6 + c) * 12 + r] = (T)a[k];
    for (int r = 0; r < 6; ++r)
      for (int c = r; c < 6; ++c, ++k) Hvb[r * 6 + c] = Hvb[c * 6 + r] = (T)a[k];
    for (int r = 0; r < 6; ++r, ++k) bvb[r] = (T)a[k];
    for (int r = 0; r < 6; ++r)
      for (int c = 0; c < 6; ++c, ++k) Hcpb[r * 6 + c] = (T)a[k];
    for (int r = 0; r < 6; ++r)
      for (int c = 0; c < 6; ++c, ++k) Hcab[r * 6 + c] = (T)a[k];
  };
  unpack(out32, d.a);
  unpack(out64, a64);
  if (st) {
    st->num_inliers = (int32_t)s.a[S_IN];
    st->num_outliers = (int32_t)s.a[S_OUT];
    st->num_bad = (int32_t)s.a[S_BAD];
    st->num_err_normal = (int32_t)s.a[S_NORM];
    st->num_err_fn = (int32_t)s.a[S_FN];
    st->pad = 0;
    st->chi2 = s.a[S_CHI2];
    st->chi2_kernelized = s.a[S_CHI2K];
    st->pert_pose_norm = st->pert_vel_norm = 0.0;
  }
  return 0;
}

// ---- tree maintenance (tree/tree_cuda_impl_.h:43-173 restated on Node_) ----

// GeometryTraits_::applyVelocities (utils/geometry_3d_impl_.h:63-72) on one
// node with its own Scalar_(_dts_mean) (Node_::applyVelocities, node_impl_.h:86-92).
static void node_apply_velocities(Node& n, const float* v) {
  const float dts = (float)n.dts_mean;
  const Vec3<float> td = mk(v[0] * dts, v[1] * dts, v[2] * dts);
  const Vec3<float> rd = mk(v[3] * dts, v[4] * dts, v[5] * dts);
  const Mat3<float> R = rodrigues(rd);
  const Vec3<float> m = add(matvec(R, mk(n.mean[0], n.mean[1], n.mean[2])), td);
  const Vec3<float> d = matvec(R, mk(n.direction[0], n.direction[1], n.direction[2]));
  for (int i = 0; i < 3; ++i) {
    n.mean[i] = m[i];
    n.direction[i] = d[i];
  }
}

// treeApplyVelocities_kernel: every node, leaves skipped when skip_leaves.
void orc_tree_apply_velocities(void* nodes, int64_t n_nodes, const float* v, int skip_leaves) {
  Node* N = (Node*)nodes;
  for (int64_t i = 0; i < n_nodes; ++i)
    if (!(N[i].isLeaf() && skip_leaves)) node_apply_velocities(N[i], v);
}

// treeApplyIsometry_kernel: applyIsometry_ on every node (tree/node_.h:127-137).
void orc_tree_apply_isometry(void* nodes, int64_t n_nodes, const float* iso) {
  Node* N = (Node*)nodes;
  const Iso3<float> X = from34<float>(iso);
  for (int64_t i = 0; i < n_nodes; ++i) {
    const Vec3<float> m = applyIsometry(X, mk(N[i].mean[0], N[i].mean[1], N[i].mean[2]));
    const Vec3<float> d = applyRotation(X, mk(N[i].direction[0], N[i].direction[1], N[i].direction[2]));
    for (int k = 0; k < 3; ++k) {
      N[i].mean[k] = m[k];
      N[i].direction[k] = d[k];
    }
  }
}

// extractCompressedLeaves(_kernel / Vel_kernel): per leaf index, mean then
// direction, warped by v when v is not NULL.
void orc_extract_leaves(const void* nodes, const uint64_t* leaves, int64_t n_leaves, const float* v, float* out) {
  const Node* N = (const Node*)nodes;
  for (int64_t i = 0; i < n_leaves; ++i) {
    Node n = N[leaves[i]];
    if (v) node_apply_velocities(n, v);
    for (int k = 0; k < 3; ++k) {
      out[6 * i + k] = n.mean[k];
      out[6 * i + 3 + k] = n.direction[k];
    }
  }
}

// ---- pieces exposed for unit checks ---------------------------------------
int orc_ldlt_solve_f(int n, const float* A, const float* rhs, float* x) {
  if (n < 1 || n > 12) return -1;
  ldlt_solve<float>(n, A, rhs, x);
  return 0;
}
int orc_ldlt_solve_d(int n, const double* A, const double* rhs, double* x) {
  if (n < 1 || n > 12) return -1;
  ldlt_solve<double>(n, A, rhs, x);
  return 0;
}

void orc_sincos(const float* x, int64_t n, float* s, float* c) {
  for (int64_t i = 0; i < n; ++i) kdh::sincos_shared(x[i], &s[i], &c[i]);
}

void orc_rodrigues_f(const float* w, float* R) {
  Mat3<float> M = rodrigues(mk(w[0], w[1], w[2]));
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) R[i * 3 + j] = M.m[i][j];
}

void orc_expmap_f(const float* v, float* X34) { to34(expmap<float>(v), X34); }

// errorAndJacobian (d3/icp_.h:41-60): returns 0, -1 (normal gate), -2 (error gate).
int orc_error_and_jacobian(const float* p_fixed, const float* n_fixed, const float* p_moving, const float* n_moving,
                           float min_normal_cos, float max_error, float* e, float* J) {
  Vec3<float> pf = mk(p_fixed[0], p_fixed[1], p_fixed[2]), nf = mk(n_fixed[0], n_fixed[1], n_fixed[2]);
  Vec3<float> pm = mk(p_moving[0], p_moving[1], p_moving[2]), nm = mk(n_moving[0], n_moving[1], n_moving[2]);
  if (dot(nf, nm) < min_normal_cos) return -1;
  Vec3<float> n = add(nm, nf);
  *e = dot(sub(pm, pf), n);
  if (fabsf(*e) > max_error) return -2;
  Vec3<float> jr = add(cross(pm, nf), cross(pf, nm));
  J[0] = n[0]; J[1] = n[1]; J[2] = n[2]; J[3] = jr[0]; J[4] = jr[1]; J[5] = jr[2];
  return 0;
}

// errorAndJacobianVelocityA (d3/cticp_impl_.h:11-39), FP32.
void orc_error_and_jacobian_velocity_a(const float* Ta34, const float* pa_local, const float* na_local,
                                       const float* pb_global, const float* nb_global, const float* va, float dt,
                                       float* e, float* J) {
  Iso3<float> Ta = from34<float>(Ta34);
  float ma[6];
  for (int i = 0; i < 6; ++i) ma[i] = va[i] * dt;
  Vec3<float> pa_m, na_m;
  Mat3<float> dRp, dRn;
  dRxp(pa_m, dRp, mk(ma[3], ma[4], ma[5]), mk(pa_local[0], pa_local[1], pa_local[2]));
  dRxp(na_m, dRn, mk(ma[3], ma[4], ma[5]), mk(na_local[0], na_local[1], na_local[2]));
  pa_m = add(pa_m, mk(ma[0], ma[1], ma[2]));
  const Mat3<float>& Ra = Ta.R;
  Vec3<float> pa_g = add(matvec(Ra, pa_m), Ta.t);
  Vec3<float> na_g = matvec(Ra, na_m);
  Vec3<float> n = add(na_g, mk(nb_global[0], nb_global[1], nb_global[2]));
  Vec3<float> dp = sub(pa_g, mk(pb_global[0], pb_global[1], pb_global[2]));
  *e = dot(dp, n);
  Mat3<float> RdRp = matmul(Ra, dRp), RdRn = matmul(Ra, dRn);
  for (int c = 0; c < 3; ++c) {
    J[c] = ((n[0] * Ra.m[0][c] + n[1] * Ra.m[1][c]) + n[2] * Ra.m[2][c]) * dt;
    float a = (n[0] * RdRp.m[0][c] + n[1] * RdRp.m[1][c]) + n[2] * RdRp.m[2][c];
    float bb = (dp[0] * RdRn.m[0][c] + dp[1] * RdRn.m[1][c]) + dp[2] * RdRn.m[2][c];
    J[3 + c] = (a + bb) * dt;
  }
}

// errorAndJacobianVelocityB (d3/cticp_impl_.h:43-73) in FP32 through the
// dual form's helper, for unit checks.
void orc_jacobian_velocity_b(const float* Tb34, const float* pa_global, const float* na_global,
                             const float* pb_local, const float* nb_local, const float* vb, float dt, float* J) {
  vel_jacobian_B<float>(J, from34<float>(Tb34), mk(pa_global[0], pa_global[1], pa_global[2]),
                        mk(na_global[0], na_global[1], na_global[2]), mk(pb_local[0], pb_local[1], pb_local[2]),
                        mk(nb_local[0], nb_local[1], nb_local[2]), vb, dt);
}

// StableAdder_<float> over a stream of scalars (for unit checks).
float orc_stable_sum_f(const float* x, int64_t n) {
  StableAdder<Packed<float, 1>> a;
  for (int64_t i = 0; i < n; ++i) {
    Packed<float, 1> p;
    p.a[0] = x[i];
    a.add(p);
  }
  return a.sum().a[0];
}

// Pose prior H/b (FP32) for unit checks against finite differences.
void orc_prior_form(const float* ZXi34, const float* X34, const float* info, float* H6, float* b6) {
  prior_form(from34<float>(ZXi34), from34<float>(X34), info, H6, b6);
}

// poseToPoseError (e, 6) and dPoseToPoseError (J, 6x6 row-major, may be NULL)
// in FP32 (geometry_3d_impl_.h:117-140).
void orc_pose_to_pose_error(const float* ZXi34, const float* X34, float* e6, float* J36) {
  pose_error_f(from34<float>(ZXi34), from34<float>(X34), e6, J36);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Tree generation: TreeGenerator_::makeTree (tree/tree_generator_impl_.h:9-19)
// restated -- the recursive Node_ constructor with pre-order indices
// (tree/node_impl_.h:29-70), LabelPolicyBase_::fillMetadata
// (tree/label_policy_base_impl_.h:13-97), CountLeafPolicy_
// (tree/count_leaf_policy_.h:27-43), FlatLeafPolicy_
// (tree/flat_leaf_policy_.h:38-108), TreeCPU_::prune / _prune,
// breadthFirstPermutation, applyPermutation and recomputeLeaves
// (tree/tree_cpu_.h:173-232, tree/tree_cpu_impl_.h:48-74, 137-158) with
// IsGoodLeafPredicate_ (tree/tree_predicates_.h:8-29).
//
// What the reference leaves to Eigen or to scheduling is fixed as the product
// builders fix it, so that their trees can be compared bit for bit:
//   * node moments: FP64 sums in the canonical order -- groups of 64
//     consecutive values combined by the pairwise tree
//     T(i, 2w) = T(i, w) + T(i + w, w) (the second term only where it
//     exists), repeated on the group sums until one value is left -- and the
//     covariance centred on the FP64 mean;
//   * eigenproblem: cyclic Jacobi sweeps in FP64 (pivots (0,1), (0,2), (1,2);
//     stop when the off-diagonal sum is below 1e-18 of the diagonal sum),
//     eigenvalues sorted ascending with ties in diagonal order;
//   * splitRange keeps the order of each side; a one-sided split makes the
//     node a leaf with its normal.
// ---------------------------------------------------------------------------
namespace orc {

struct TreeGenParams {  // kdh_tree_gen_params layout
  int32_t propagate_normal, tree_min_good_leaves, leaf_min_points, leaf_max_points;
  float leaf_density, leaf_min_normal_incidence, leaf_eigenvalue_threshold, leaf_normal_eigenratio;
  int32_t parallel_level;
  int32_t label_mode;  // 1: the reference's FP32 arithmetic (below)
};

// SelfAdjointEigenSolver<Matrix3f>::computeDirect (Eigen 3.4, which the
// reference takes from the system: kd_slam2/CMakeLists.txt:26), from its
// published closed form.  Matrix shifted by trace / 3 and divided by its
// largest |entry|; eigenvalues from the trigonometric solution of the
// characteristic cubic; eigenvectors from cross products of the columns of
// the rank-2 matrices A - lambda I.  Parity with Eigen's own bits unpinned:
// Eigen is absent here.
struct EigDirect {
  float w[3];          // ascending
  Vec3<float> v[3];    // v[k]: eigenvector of w[k]
};

static Vec3<float> scaled(Vec3<float> a, float f) { return mk(a[0] * f, a[1] * f, a[2] * f); }
static Vec3<float> divided(Vec3<float> a, float d) { return mk(a[0] / d, a[1] / d, a[2] / d); }

// extract_kernel: the column of largest |diagonal| (first on ties) crossed
// with each following column; the product of larger squared norm, unit.
static void kernel_of(const Mat3<float>& M, Vec3<float>& res, Vec3<float>& representative) {
  int i0 = 0;
  for (int i = 1; i < 3; ++i)
    if (fabsf(M.m[i][i]) > fabsf(M.m[i0][i0])) i0 = i;
  auto col = [&](int c) { return mk(M.m[0][c], M.m[1][c], M.m[2][c]); };
  representative = col(i0);
  const Vec3<float> a = cross(representative, col((i0 + 1) % 3));
  const Vec3<float> b = cross(representative, col((i0 + 2) % 3));
  const float na = dot(a, a), nb = dot(b, b);
  res = na > nb ? divided(a, sqrtf(na)) : divided(b, sqrtf(nb));
}

static EigDirect eig_direct(const Mat3<float>& A) {
  EigDirect out;
  const float third = 1.0f / 3.0f;
  const float shift = ((A.m[0][0] + A.m[1][1]) + A.m[2][2]) / 3.0f;
  Mat3<float> M;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) M.m[r][c] = A.m[r < c ? c : r][r < c ? r : c];
  for (int i = 0; i < 3; ++i) M.m[i][i] -= shift;
  float scale = 0.0f;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) scale = std::max(scale, fabsf(M.m[r][c]));
  if (scale > 0.0f)
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) M.m[r][c] = M.m[r][c] / scale;
  const float a00 = M.m[0][0], a11 = M.m[1][1], a22 = M.m[2][2];
  const float a10 = M.m[1][0], a20 = M.m[2][0], a21 = M.m[2][1];
  // det, the sum of the principal 2x2 minors and the trace
  const float det = a00 * a11 * a22 + 2.0f * a10 * a20 * a21 - a00 * a21 * a21 - a11 * a20 * a20 - a22 * a10 * a10;
  const float minors = a00 * a11 - a10 * a10 + a00 * a22 - a20 * a20 + a11 * a22 - a21 * a21;
  const float tr = a00 + a11 + a22;
  const float m = tr * third;
  const float p = std::max((tr * m - minors) * third, 0.0f);
  const float hb = 0.5f * (det + m * (2.0f * m * m - minors));
  const float disc = std::max(p * p * p - hb * hb, 0.0f);
  const float r = sqrtf(p);
  const float phi = atan2f(sqrtf(disc), hb) * third;
  const float cph = cosf(phi), sph = sinf(phi), s3 = sqrtf(3.0f);
  out.w[0] = m - r * (cph + s3 * sph);
  out.w[1] = m - r * (cph - s3 * sph);
  out.w[2] = m + 2.0f * r * cph;
  const float eps = std::numeric_limits<float>::epsilon();
  if (out.w[2] - out.w[0] <= eps) {
    out.v[0] = mk(1.0f, 0.0f, 0.0f);
    out.v[1] = mk(0.0f, 1.0f, 0.0f);
    out.v[2] = mk(0.0f, 0.0f, 1.0f);
  } else {
    const float gap_hi = out.w[2] - out.w[1], gap_lo = out.w[1] - out.w[0];
    const bool top_first = gap_hi > gap_lo;  // the most separated eigenvalue is the largest
    const int k = top_first ? 2 : 0, l = top_first ? 0 : 2;
    const float dk = top_first ? gap_lo : gap_hi;  // the smaller gap
    const float dl = gap_lo;
    Mat3<float> T = M;
    for (int i = 0; i < 3; ++i) T.m[i][i] -= out.w[k];
    kernel_of(T, out.v[k], out.v[l]);
    if (dk <= 2.0f * eps * dl) {
      const float pr = dot(out.v[k], out.v[l]);
      out.v[l] = sub(out.v[l], scaled(out.v[l], pr));
      const float z = dot(out.v[l], out.v[l]);
      if (z > 0.0f) out.v[l] = divided(out.v[l], sqrtf(z));
    } else {
      T = M;
      for (int i = 0; i < 3; ++i) T.m[i][i] -= out.w[l];
      Vec3<float> ignored;
      kernel_of(T, out.v[l], ignored);
    }
    const Vec3<float> mid = cross(out.v[2], out.v[0]);
    const float z = dot(mid, mid);
    out.v[1] = z > 0.0f ? divided(mid, sqrtf(z)) : mid;
  }
  for (int k = 0; k < 3; ++k) out.w[k] = out.w[k] * scale + shift;
  return out;
}

struct TPoint {
  float c[3];
  double t;
};

template <int K>
using Acc = std::array<double, K>;

// T(i, w) over v[lo, lo + m): the pairwise tree of one group
template <int K>
static Acc<K> pair_tree(const std::vector<Acc<K>>& v, size_t lo, size_t m, size_t i, size_t w) {
  if (w == 1) return v[lo + i];
  Acc<K> a = pair_tree<K>(v, lo, m, i, w / 2);
  if (i + w / 2 < m) {
    const Acc<K> b = pair_tree<K>(v, lo, m, i + w / 2, w / 2);
    for (int k = 0; k < K; ++k) a[k] += b[k];
  }
  return a;
}

template <int K, typename F>
static Acc<K> canonical_sum(size_t s, size_t e, F item) {
  std::vector<Acc<K>> v;
  for (size_t i = s; i < e; ++i) v.push_back(item(i));
  while (v.size() > 1) {
    std::vector<Acc<K>> w;
    for (size_t lo = 0; lo < v.size(); lo += 64) w.push_back(pair_tree<K>(v, lo, std::min<size_t>(64, v.size() - lo), 0, 64));
    v.swap(w);
  }
  return v[0];
}

static void jacobi3(double A[3][3], double lam[3], double V[3][3]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) V[i][j] = i == j;
  static const int PQ[3][2] = {{0, 1}, {0, 2}, {1, 2}};
  for (int sweep = 0; sweep < 50; ++sweep) {
    const double off = (fabs(A[0][1]) + fabs(A[0][2])) + fabs(A[1][2]);
    const double dia = (fabs(A[0][0]) + fabs(A[1][1])) + fabs(A[2][2]);
    if (off <= 1e-300 || off <= 1e-18 * dia) break;
    for (const auto& pq : PQ) {
      const int p = pq[0], q = pq[1];
      if (fabs(A[p][q]) <= 1e-300) continue;
      const double th = (A[q][q] - A[p][p]) / (2.0 * A[p][q]);
      const double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
      const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
      for (int k = 0; k < 3; ++k) {
        const double x = A[k][p], y = A[k][q];
        A[k][p] = c * x - s * y;
        A[k][q] = s * x + c * y;
      }
      for (int k = 0; k < 3; ++k) {
        const double x = A[p][k], y = A[q][k];
        A[p][k] = c * x - s * y;
        A[q][k] = s * x + c * y;
      }
      for (int k = 0; k < 3; ++k) {
        const double x = V[k][p], y = V[k][q];
        V[k][p] = c * x - s * y;
        V[k][q] = s * x + c * y;
      }
    }
  }
  int order[3] = {0, 1, 2};
  std::stable_sort(order, order + 3, [&](int a, int b) { return A[a][a] < A[b][b]; });
  double W[3][3];
  for (int k = 0; k < 3; ++k) {
    lam[k] = A[order[k]][order[k]];
    for (int i = 0; i < 3; ++i) W[i][k] = V[i][order[k]];
  }
  memcpy(V, W, sizeof(W));
}

// LabelPolicyMetadata of the three policies
struct TMeta {
  size_t n_points = 0;
  int level = 0;
  double ts_start = 0;
  Vec3<float> mean{};
  double lam[3] = {0, 0, 0};
  Vec3<float> evec[3]{};  // column k: k-th eigenvector
  bool is_bad = true, is_node = false, is_leaf = false, is_flat = false, has_propagated_normal = false;
  Vec3<float> propagated_normal{};
};

struct TreeGen {
  TreeGenParams P;
  TPoint* pts;
  std::vector<Node> nodes;
  int idx = 0;

  static Vec3<float> dir_of(const Node& n) { return mk(n.direction[0], n.direction[1], n.direction[2]); }
  static void set_dir(Node& n, Vec3<float> d) {
    for (int i = 0; i < 3; ++i) n.direction[i] = d[i];
  }

  // LabelPolicyBase_::fillMetadata with the reference's arithmetic
  // (label_policy_base_impl_.h:41-95): FP32 raw moments through StableAdder_,
  // E[x x^T] - mu mu^T, stamp moments through a double StableAdder_ scaled by
  // the FP32 1/n, computeDirect.
  void fill_base_ref(TMeta& m, Node& node) {
    const float isize = 1.0f / (float)m.n_points;
    StableAdder<Packed<double, 2>> ts_adder;
    StableAdder<Packed<float, 9>> acc;
    for (size_t i = node.idx_start; i < node.idx_end; ++i) {
      const double d = pts[i].t - m.ts_start;
      Packed<double, 2> tv;
      tv.a[0] = d;
      tv.a[1] = d * d;
      ts_adder.add(tv);
      Packed<float, 9> v;
      const float* c = pts[i].c;
      int k = 0;
      for (int r = 0; r < 3; ++r)
        for (int q = r; q < 3; ++q) v.a[k++] = c[r] * c[q];
      for (int r = 0; r < 3; ++r) v.a[k++] = c[r];
      acc.add(v);
    }
    const Packed<double, 2> ts2 = ts_adder.sum();
    const double t0 = ts2.a[0] * (double)isize, t1 = ts2.a[1] * (double)isize;
    node.dts_mean = t0;
    node.dts_covariance = t1 - t0 * t0;
    Packed<float, 9> sum = acc.sum();
    for (int k = 0; k < 9; ++k) sum.a[k] = sum.a[k] * isize;
    Mat3<float> cov;
    int k = 0;
    for (int r = 0; r < 3; ++r)
      for (int q = r; q < 3; ++q, ++k) cov.m[r][q] = cov.m[q][r] = sum.a[k];
    m.mean = mk(sum.a[6], sum.a[7], sum.a[8]);
    for (int r = 0; r < 3; ++r) node.mean[r] = m.mean[r];
    for (int r = 0; r < 3; ++r)
      for (int q = 0; q < 3; ++q) cov.m[r][q] = cov.m[r][q] - m.mean[r] * m.mean[q];
    if (m.n_points < 3) {
      set_dir(node, mk(0.0f, 0.0f, 0.0f));
      m.is_bad = true;
      return;
    }
    const EigDirect e = eig_direct(cov);
    for (int q = 0; q < 3; ++q) {
      m.lam[q] = (double)e.w[q];
      m.evec[q] = e.v[q];
    }
    if (dot(m.mean, m.evec[0]) > 0) m.evec[0] = mk(-m.evec[0][0], -m.evec[0][1], -m.evec[0][2]);
    if (dot(m.mean, m.evec[2]) > 0) m.evec[2] = mk(-m.evec[2][0], -m.evec[2][1], -m.evec[2][2]);
    m.is_bad = false;
  }

  void fill_base(TMeta& m, Node& node, const TMeta* parent) {
    m.n_points = node.idx_end - node.idx_start;
    if (!m.n_points) {
      m.is_bad = true;
      return;
    }
    if (!parent) {
      double ts = std::numeric_limits<double>::max();
      for (size_t i = node.idx_start; i < node.idx_end; ++i) ts = std::min(ts, pts[i].t);
      m.ts_start = ts;
    } else {
      m.ts_start = parent->ts_start;
    }
    m.level = parent ? parent->level + 1 : 0;
    if (P.label_mode == 1) {
      fill_base_ref(m, node);
      return;
    }
    const double isize = 1.0 / (double)m.n_points;
    const Acc<5> s1 = canonical_sum<5>(node.idx_start, node.idx_end, [&](size_t i) {
      const double d = pts[i].t - m.ts_start;
      return Acc<5>{(double)pts[i].c[0], (double)pts[i].c[1], (double)pts[i].c[2], d, d * d};
    });
    const double ts1 = s1[3] * isize, ts2 = s1[4] * isize;
    node.dts_mean = ts1;
    node.dts_covariance = ts2 - ts1 * ts1;
    const double mu[3] = {s1[0] * isize, s1[1] * isize, s1[2] * isize};
    for (int k = 0; k < 3; ++k) node.mean[k] = m.mean[k] = (float)mu[k];
    if (m.n_points < 3) {
      set_dir(node, mk(0.0f, 0.0f, 0.0f));
      m.is_bad = true;
      return;
    }
    const Acc<6> c2 = canonical_sum<6>(node.idx_start, node.idx_end, [&](size_t i) {
      const double a = (double)pts[i].c[0] - mu[0], b = (double)pts[i].c[1] - mu[1], c = (double)pts[i].c[2] - mu[2];
      return Acc<6>{a * a, a * b, a * c, b * b, b * c, c * c};
    });
    double C[3][3] = {{c2[0] * isize, c2[1] * isize, c2[2] * isize},
                      {c2[1] * isize, c2[3] * isize, c2[4] * isize},
                      {c2[2] * isize, c2[4] * isize, c2[5] * isize}};
    double V[3][3];
    jacobi3(C, m.lam, V);
    for (int k = 0; k < 3; ++k) m.evec[k] = mk((float)V[0][k], (float)V[1][k], (float)V[2][k]);
    if (dot(m.mean, m.evec[0]) > 0) m.evec[0] = mk(-m.evec[0][0], -m.evec[0][1], -m.evec[0][2]);
    if (dot(m.mean, m.evec[2]) > 0) m.evec[2] = mk(-m.evec[2][0], -m.evec[2][1], -m.evec[2][2]);
    m.is_bad = false;
  }

  bool propagate_normal(Node& node, TMeta& m) {
    m.is_leaf = true;
    if (m.has_propagated_normal) {
      set_dir(node, m.propagated_normal);
      if (dot(m.mean, dir_of(node)) > 0) set_dir(node, mk(-node.direction[0], -node.direction[1], -node.direction[2]));
      return true;
    }
    return false;
  }

  void fill(TMeta& m, Node& node, const TMeta* parent) {
    fill_base(m, node, parent);
    // CountLeafPolicy_
    if (m.is_bad) {
      m.is_leaf = true;
    } else if ((double)m.n_points > (double)P.leaf_max_points) {
      m.is_node = true;
      set_dir(node, m.evec[2]);
    } else {
      m.is_leaf = true;
      set_dir(node, m.evec[0]);
    }
    // FlatLeafPolicy_
    if (parent && P.propagate_normal) {
      m.propagated_normal = parent->propagated_normal;
      m.has_propagated_normal = parent->has_propagated_normal;
    } else {
      m.propagated_normal = mk(0.0f, 0.0f, 0.0f);
      m.has_propagated_normal = false;
    }
    if (m.is_bad || (double)m.n_points < (double)P.leaf_min_points) {
      m.is_bad = !propagate_normal(node, m);
      return;
    }
    const float lmax = (float)m.lam[2], lmin = (float)m.lam[0];
    m.is_flat = lmin / lmax < P.leaf_normal_eigenratio && fabsf(dot(m.mean, m.evec[0])) > P.leaf_min_normal_incidence;
    if (m.is_flat) {
      m.has_propagated_normal = true;
      m.propagated_normal = m.evec[0];
    }
    if (m.is_node) return;
    float volume = 1.0f;
    for (int i = 1; i < 3; ++i) volume *= (float)m.lam[i];
    if ((float)m.n_points / volume < P.leaf_density) {
      m.is_leaf = false;
      m.is_node = true;
      m.is_bad = false;
      set_dir(node, m.evec[2]);
      return;
    }
    if (m.is_leaf && m.is_flat) {
      set_dir(node, m.evec[0]);
      return;
    }
    m.is_bad = !propagate_normal(node, m);
  }

  // Node_ constructor: label, split, children at pre-order indices
  void make(int me, size_t s, size_t e, const TMeta* parent) {
    Node node;
    memset(&node, 0, sizeof(node));
    node.idx_start = s;
    node.idx_end = e;
    node.idx_left = node.idx_right = -1;
    TMeta m;
    fill(m, node, parent);
    if (m.is_node) {
      auto left_side = [&](const TPoint& q) {
        const Vec3<float> d = mk(q.c[0] - node.mean[0], q.c[1] - node.mean[1], q.c[2] - node.mean[2]);
        return dot(dir_of(node), d) > 0;
      };
      size_t mid;
      if (P.label_mode == 1) {
        // splitRange (node_impl_.h:12-26): right-side points swapped to the
        // back; signed bounds (the reference's unsigned end wraps below 0)
        long lo = (long)s, hi = (long)e - 1;
        while (lo <= hi) {
          if (left_side(pts[lo])) {
            ++lo;
          } else {
            std::swap(pts[lo], pts[hi]);
            --hi;
          }
        }
        mid = (size_t)lo;
      } else {
        // splitRange, each side kept in order
        std::vector<TPoint> left, right;
        for (size_t i = s; i < e; ++i) (left_side(pts[i]) ? left : right).push_back(pts[i]);
        mid = s + left.size();
        std::copy(left.begin(), left.end(), pts + s);
        std::copy(right.begin(), right.end(), pts + mid);
      }
      if (mid == s || mid == e) {
        set_dir(node, m.evec[0]);
      } else {
        nodes[me] = node;
        node.idx_left = ++idx;
        make(node.idx_left, s, mid, &m);
        node.idx_right = ++idx;
        make(node.idx_right, mid, e, &m);
      }
    }
    nodes[me] = node;
  }

  static bool is_leaf(const Node& n) { return n.idx_left < 0 && n.idx_right < 0; }

  int prune(int i) {
    Node& n = nodes[i];
    if (is_leaf(n)) {
      const Vec3<float> d = dir_of(n), mean = mk(n.mean[0], n.mean[1], n.mean[2]);
      const bool good = dot(d, d) > 0 && fabsf(dot(mean, d)) > 0.1f;
      return good ? i : -1;
    }
    if (n.idx_left >= 0) n.idx_left = prune(n.idx_left);
    if (n.idx_right >= 0) n.idx_right = prune(n.idx_right);
    if (is_leaf(n)) {
      set_dir(n, mk(0.0f, 0.0f, 0.0f));
      return -1;
    }
    return i;
  }
};

}  // namespace orc

extern "C" {

// computeDirect of a symmetric 3x3 (row-major A, lower triangle read):
// eigenvalues ascending in w, eigenvectors as the columns of V (row-major).
void orc_eig_direct(const float* A9, float* w, float* V9) {
  using namespace orc;
  Mat3<float> A;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) A.m[r][c] = A9[3 * r + c];
  const EigDirect e = eig_direct(A);
  for (int k = 0; k < 3; ++k) {
    w[k] = e.w[k];
    for (int r = 0; r < 3; ++r) V9[3 * r + k] = e.v[k][r];
  }
}

// makeTree on n points (xyz 3 floats each, stamps or NULL).  Returns 1 and
// fills the BFS-ordered nodes (capacity 2n), the leaf list (capacity n) and
// the permuted points / stamps; returns 0 where makeTree returns nullptr
// (fewer than leaf_min_points points, fewer than tree_min_good_leaves leaves,
// or no points).
int orc_tree_build(const float* xyz, const double* stamps, int64_t n, const void* params, void* nodes_out,
                   int64_t* n_nodes, uint64_t* leaves_out, int64_t* n_leaves, float* xyz_out, double* stamps_out) {
  using namespace orc;
  *n_nodes = *n_leaves = 0;
  TreeGen g;
  memcpy(&g.P, params, sizeof(TreeGenParams));
  if (n < (int64_t)g.P.leaf_min_points || n == 0) return 0;
  std::vector<TPoint> pts((size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    for (int k = 0; k < 3; ++k) pts[i].c[k] = xyz[3 * i + k];
    pts[i].t = stamps ? stamps[i] : 0.0;
  }
  g.pts = pts.data();
  g.nodes.assign((size_t)(2 * n), Node{});
  g.make(0, 0, (size_t)n, nullptr);
  g.nodes.resize((size_t)g.idx + 1);
  std::vector<Node> out;
  if (g.prune(0) >= 0) {
    // breadthFirstPermutation + applyPermutation
    std::vector<int> perm, inv(g.nodes.size(), -1);
    std::deque<int> q{0};
    while (!q.empty()) {
      const int c = q.front();
      q.pop_front();
      perm.push_back(c);
      if (g.nodes[c].idx_left >= 0) q.push_back(g.nodes[c].idx_left);
      if (g.nodes[c].idx_right >= 0) q.push_back(g.nodes[c].idx_right);
    }
    for (size_t i = 0; i < perm.size(); ++i) inv[perm[i]] = (int)i;
    for (int p : perm) {
      Node d = g.nodes[p];
      if (d.idx_left >= 0) d.idx_left = inv[d.idx_left];
      if (d.idx_right >= 0) d.idx_right = inv[d.idx_right];
      out.push_back(d);
    }
  }
  std::vector<uint64_t> leaves;  // recomputeLeaves
  for (size_t i = 0; i < out.size(); ++i)
    if (TreeGen::is_leaf(out[i])) leaves.push_back(i);
  if ((int64_t)leaves.size() < (int64_t)g.P.tree_min_good_leaves) return 0;
  memcpy(nodes_out, out.data(), out.size() * sizeof(Node));
  memcpy(leaves_out, leaves.data(), leaves.size() * sizeof(uint64_t));
  *n_nodes = (int64_t)out.size();
  *n_leaves = (int64_t)leaves.size();
  for (int64_t i = 0; i < n; ++i) {
    for (int k = 0; k < 3; ++k) xyz_out[3 * i + k] = pts[i].c[k];
    if (stamps_out) stamps_out[i] = pts[i].t;
  }
  return 1;
}

}  // extern "C"
