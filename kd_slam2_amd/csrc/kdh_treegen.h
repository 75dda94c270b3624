// PCA KD-tree generation: the parts shared by the host builder
// (tree_build.cpp) and the device builder (kdh_tree_build.hip), so that both
// produce the same tree bit for bit.
//
// Follows TreeGenerator_::makeTree (kd_slam2/src/lib/kd_slam/tree/
// tree_generator_impl_.h:9-19): Node_ construction and splitRange
// (tree/node_impl_.h:13-70), LabelPolicyBase_::fillMetadata
// (tree/label_policy_base_impl_.h:13-97), CountLeafPolicy_
// (tree/count_leaf_policy_.h:27-43), FlatLeafPolicy_
// (tree/flat_leaf_policy_.h:38-108) and IsGoodLeafPredicate_
// (tree/tree_predicates_.h:8-29).  Where this build fixes what the reference
// leaves to Eigen or to thread scheduling:
//   * node moments are FP64 sums in one canonical order (below), the
//     covariance centred on the FP64 mean (the reference sums E[x x^T] in FP32
//     and subtracts mu mu^T, which cancels badly far from the origin);
//   * the 3x3 eigenproblem is cyclic Jacobi in FP64 with a stable ascending
//     sort (the reference calls SelfAdjointEigenSolver::computeDirect);
//   * a split keeps the relative order of the points on each side (the
//     reference's splitRange swaps from both ends): the point order inside a
//     node, and with it every sum, is then a function of the input order only;
//   * a split that leaves one side empty makes the node a leaf with its
//     normal (the reference would recurse on the same range without end).
//
// Canonical summation order of a node's n items (points in node order):
//   while more than one value is left, groups of 64 consecutive values are
//   combined by a stride-doubling pairwise tree (for s = 1, 2, .., 32:
//   v[i] += v[i + s] for i = 0 mod 2s, i + s < group size) into one value per
//   group.  The first round turns the points of each 64-point chunk into one
//   chunk sum.  A wave evaluates one group exactly: lane i holds value i and
//   the tree is six rounds of shuffles.
#pragma once

#include "../../include/kdh_host.h"
#include "kdh_math.h"

#pragma clang fp contract(off)

namespace kdh {
namespace tg {

constexpr int kChunk = 64;
constexpr int kGroup = 64;

// One group of m <= 64 values combined in place into v[0] (host form).
template <int K>
KDH_HD void group_pairwise(double (*v)[K], int m) {
  for (int s = 1; s < m; s <<= 1)
    for (int i = 0; i + s < m; i += 2 * s)
      for (int k = 0; k < K; ++k) v[i][k] += v[i + s][k];
}

// The canonical reduction of m values (host form); v is overwritten and the
// result is left in v[0].
template <int K>
KDH_HD void reduce_canonical(double (*v)[K], long m) {
  while (m > 1) {
    const long groups = (m + kGroup - 1) / kGroup;
    for (long g = 0; g < groups; ++g) {
      const long b = g * kGroup;
      const int cnt = (int)((m - b) < kGroup ? (m - b) : kGroup);
      group_pairwise<K>(v + b, cnt);
      if (g > 0)
        for (int k = 0; k < K; ++k) v[g][k] = v[b][k];
    }
    m = groups;
  }
}

KDH_HD float dot3f(const float a[3], const float b[3]) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

// Cyclic Jacobi eigen-decomposition of a symmetric 3x3 matrix (FP64):
// eigenvalues ascending in w, unit eigenvectors as columns of V.  Ties keep
// the diagonal order (stable insertion sort).
KDH_HD void eig3_jacobi(const double A_in[3][3], double w[3], double V[3][3]) {
  double A[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      A[i][j] = A_in[i][j];
      V[i][j] = (i == j) ? 1.0 : 0.0;
    }
  for (int sweep = 0; sweep < 50; ++sweep) {
    const double off = (fabs(A[0][1]) + fabs(A[0][2])) + fabs(A[1][2]);
    const double scale = (fabs(A[0][0]) + fabs(A[1][1])) + fabs(A[2][2]);
    if (off <= 1e-300 || off <= 1e-18 * scale) break;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        if (fabs(A[p][q]) <= 1e-300) continue;
        const double theta = (A[q][q] - A[p][p]) / (2.0 * A[p][q]);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < 3; ++k) {  // A <- A J
          const double akp = A[k][p], akq = A[k][q];
          A[k][p] = c * akp - s * akq;
          A[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < 3; ++k) {  // A <- J^T A
          const double apk = A[p][k], aqk = A[q][k];
          A[p][k] = c * apk - s * aqk;
          A[q][k] = s * apk + c * aqk;
        }
        for (int k = 0; k < 3; ++k) {
          const double vkp = V[k][p], vkq = V[k][q];
          V[k][p] = c * vkp - s * vkq;
          V[k][q] = s * vkp + c * vkq;
        }
      }
  }
  int idx[3] = {0, 1, 2};
  const double d[3] = {A[0][0], A[1][1], A[2][2]};
  for (int i = 1; i < 3; ++i)
    for (int j = i; j > 0 && d[idx[j]] < d[idx[j - 1]]; --j) {
      const int t = idx[j];
      idx[j] = idx[j - 1];
      idx[j - 1] = t;
    }
  double Vs[3][3];
  for (int k = 0; k < 3; ++k) {
    w[k] = d[idx[k]];
    for (int i = 0; i < 3; ++i) Vs[i][k] = V[i][idx[k]];
  }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) V[i][j] = Vs[i][j];
}

// What the labelling of one node decides.
struct Label {
  float mean[3];
  float dir[3];
  float evec0[3];      // the normal, used when a split turns out one-sided
  float prop[3];       // FlatLeafPolicy_ propagated normal handed to the children
  int has_prop;
  int is_node;         // split this node
  double dts_mean, dts_cov;
};

// CountLeafPolicy_ and FlatLeafPolicy_ after LabelPolicyBase_::fillMetadata:
// n points, is_bad (fewer than 3 points), the node mean, the eigenvalues
// (ascending) and the unit eigenvectors evec[k] (the normal evec[0] and the
// direction evec[2] already reoriented against the mean).  L.mean, dts_mean
// and dts_cov are set by the caller.  inherited prop / has_prop: the
// parent's (ignored for the root or without propagate_normal).
KDH_HD void label_policies(const kdh_tree_gen_params& P, long n, bool is_bad, const double evals[3],
                           const float evec[3][3], bool root, const float inherited_prop[3], int inherited_has,
                           Label& L) {
  bool is_leaf = false, is_node = false;
  for (int k = 0; k < 3; ++k) L.dir[k] = L.evec0[k] = 0.0f;
  if (!is_bad)
    for (int i = 0; i < 3; ++i) L.evec0[i] = evec[0][i];
  // CountLeafPolicy_
  if (is_bad) {
    is_leaf = true;
  } else if ((double)n > (double)P.leaf_max_points) {
    is_node = true;
    for (int i = 0; i < 3; ++i) L.dir[i] = evec[2][i];
  } else {
    is_leaf = true;
    for (int i = 0; i < 3; ++i) L.dir[i] = evec[0][i];
  }
  // FlatLeafPolicy_
  bool has_prop = false;
  float prop[3] = {0.0f, 0.0f, 0.0f};
  if (!root && P.propagate_normal) {
    has_prop = inherited_has != 0;
    for (int i = 0; i < 3; ++i) prop[i] = inherited_prop[i];
  }
  auto propagate = [&]() {  // propagateNormal: marks the node a leaf
    is_leaf = true;
    if (has_prop) {
      for (int i = 0; i < 3; ++i) L.dir[i] = prop[i];
      if (dot3f(L.mean, L.dir) > 0)
        for (int i = 0; i < 3; ++i) L.dir[i] = -L.dir[i];
    }
  };
  bool done = false;
  if (is_bad || (double)n < (double)P.leaf_min_points) {
    propagate();
    done = true;
  }
  if (!done) {
    const float lmax = (float)evals[2], lmin = (float)evals[0];
    const float ratio = lmin / lmax;
    const bool is_flat = ratio < P.leaf_normal_eigenratio && fabsf(dot3f(L.mean, evec[0])) > P.leaf_min_normal_incidence;
    if (is_flat) {
      has_prop = true;
      for (int i = 0; i < 3; ++i) prop[i] = evec[0][i];
    }
    if (!is_node) {
      float volume = 1.0f;
      for (int i = 1; i < 3; ++i) volume *= (float)evals[i];
      const float density = (float)n / volume;
      if (density < P.leaf_density) {
        is_leaf = false;
        is_node = true;
        for (int i = 0; i < 3; ++i) L.dir[i] = evec[2][i];
      } else if (is_leaf && is_flat) {
        for (int i = 0; i < 3; ++i) L.dir[i] = evec[0][i];
      } else {
        propagate();
      }
    }
  }
  L.is_node = is_node ? 1 : 0;
  L.has_prop = has_prop ? 1 : 0;
  for (int i = 0; i < 3; ++i) L.prop[i] = prop[i];
}

// LabelPolicyBase_::fillMetadata after the sums, then the policies.  n:
// points; s1 = (sum x, sum y, sum z, sum dts, sum dts^2) in the canonical
// order; c2 = centred sums (xx, xy, xz, yy, yz, zz) about the FP64 mean
// s1[0..2] / n in the canonical order.
KDH_HD void label_node(const kdh_tree_gen_params& P, long n, const double s1[5], const double c2[6], bool root,
                       const float inherited_prop[3], int inherited_has, Label& L) {
  const double inv = 1.0 / (double)n;
  const double t1 = s1[3] * inv, t2 = s1[4] * inv;
  L.dts_mean = t1;
  L.dts_cov = t2 - t1 * t1;
  for (int k = 0; k < 3; ++k) L.mean[k] = (float)(s1[k] * inv);
  bool is_bad = true;
  float evec[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  double evals[3] = {0, 0, 0};
  if (n >= 3) {
    double C[3][3];
    C[0][0] = c2[0] * inv;
    C[0][1] = C[1][0] = c2[1] * inv;
    C[0][2] = C[2][0] = c2[2] * inv;
    C[1][1] = c2[3] * inv;
    C[1][2] = C[2][1] = c2[4] * inv;
    C[2][2] = c2[5] * inv;
    double V[3][3];
    eig3_jacobi(C, evals, V);
    for (int k = 0; k < 3; ++k)
      for (int i = 0; i < 3; ++i) evec[k][i] = (float)V[i][k];
    // reorient the normal (smallest) and the direction (largest) against the mean
    if (dot3f(L.mean, evec[0]) > 0)
      for (int i = 0; i < 3; ++i) evec[0][i] = -evec[0][i];
    if (dot3f(L.mean, evec[2]) > 0)
      for (int i = 0; i < 3; ++i) evec[2][i] = -evec[2][i];
    is_bad = false;
  }
  label_policies(P, n, is_bad, evals, evec, root, inherited_prop, inherited_has, L);
}

// ---- label_mode 1: the reference's own arithmetic (host builder only) ------
//
// LabelPolicyBase_::fillMetadata as the reference evaluates it
// (tree/label_policy_base_impl_.h:41-95): raw FP32 moments E[x x^T], E[x]
// summed by StableAdder_ in point order, the covariance E[x x^T] - mu mu^T in
// FP32, stamp moments by a double StableAdder_ scaled by the FP32 1 / n, and
// the eigenproblem by SelfAdjointEigenSolver<Matrix3f>::computeDirect.  Eigen
// is not vendored by the reference (kd_slam2/CMakeLists.txt:26,
// find_package(Eigen3), 3.4 on its ROS 2 Humble base); eig3_direct restates
// that published closed form.  Three-term sums (traces, squared norms, dot
// products) are evaluated left to right, as everywhere in this build.

// StableAdder_ (utils/stable_adder_.h:21-74): a value goes into level 0; a
// level that has taken 1024 values passes its sum on to the next level and
// starts again from zero; the total is the five levels added in order.
template <typename T, int K>
struct LevelAdder {
  static constexpr int kLevels = 5, kBatch = 1024;
  T v[kLevels][K];
  int n[kLevels];
  LevelAdder() {
    for (int i = 0; i < kLevels; ++i) {
      n[i] = 0;
      for (int k = 0; k < K; ++k) v[i][k] = T(0);
    }
  }
  void add(const T* x) {
    T cur[K];
    for (int k = 0; k < K; ++k) cur[k] = x[k];
    for (int i = 0; i < kLevels; ++i) {
      for (int k = 0; k < K; ++k) v[i][k] += cur[k];
      if (++n[i] < kBatch) return;
      for (int k = 0; k < K; ++k) {
        cur[k] = v[i][k];
        v[i][k] = T(0);
      }
      n[i] = 0;
    }
  }
  void sum(T* out) const {
    for (int k = 0; k < K; ++k) out[k] = T(0);
    for (int i = 0; i < kLevels; ++i)
      for (int k = 0; k < K; ++k) out[k] += v[i][k];
  }
};

inline void cross3f(const float a[3], const float b[3], float c[3]) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

// A unit vector orthogonal to the two most independent columns of the
// rank-2 matrix M (column i0 of the largest |diagonal| crossed with each of
// the others, the longer product normalised); rep gets column i0.
inline void kernel_vector(const float M[3][3], float out[3], float rep[3]) {
  int i0 = 0;
  float best = fabsf(M[0][0]);
  for (int i = 1; i < 3; ++i)
    if (fabsf(M[i][i]) > best) {
      best = fabsf(M[i][i]);
      i0 = i;
    }
  float ci[3], c0[3], c1[3];
  for (int r = 0; r < 3; ++r) ci[r] = M[r][i0];
  const int j0 = (i0 + 1) % 3, j1 = (i0 + 2) % 3;
  const float cj0[3] = {M[0][j0], M[1][j0], M[2][j0]};
  const float cj1[3] = {M[0][j1], M[1][j1], M[2][j1]};
  cross3f(ci, cj0, c0);
  cross3f(ci, cj1, c1);
  const float n0 = dot3f(c0, c0), n1 = dot3f(c1, c1);
  if (n0 > n1) {
    const float d = sqrtf(n0);
    for (int r = 0; r < 3; ++r) out[r] = c0[r] / d;
  } else {
    const float d = sqrtf(n1);
    for (int r = 0; r < 3; ++r) out[r] = c1[r] / d;
  }
  for (int r = 0; r < 3; ++r) rep[r] = ci[r];
}

// SelfAdjointEigenSolver<Matrix3f>::computeDirect: the matrix shifted by its
// mean eigenvalue and scaled into [-1, 1]; the eigenvalues as the roots of
// the characteristic polynomial in trigonometric form (ascending); the
// eigenvector of the most separated eigenvalue as the kernel of A - lambda I,
// the second likewise (or orthogonalised when the other two coincide), the
// middle one as their cross product.  A symmetric; w ascending; V columns.
inline void eig3_direct(const float A[3][3], float w[3], float V[3][3]) {
  const float inv3 = 1.0f / 3.0f;
  const float sqrt3 = sqrtf(3.0f);
  const float shift = ((A[0][0] + A[1][1]) + A[2][2]) / 3.0f;
  float S[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) S[r][c] = r >= c ? A[r][c] : A[c][r];  // the lower triangle, mirrored
  for (int i = 0; i < 3; ++i) S[i][i] -= shift;
  float scale = 0.0f;
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) scale = fabsf(S[r][c]) > scale ? fabsf(S[r][c]) : scale;
  if (scale > 0.0f)
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) S[r][c] /= scale;
  // roots of x^3 - c2 x^2 + c1 x - c0
  const float m00 = S[0][0], m11 = S[1][1], m22 = S[2][2], m10 = S[1][0], m20 = S[2][0], m21 = S[2][1];
  const float c0 = m00 * m11 * m22 + 2.0f * m10 * m20 * m21 - m00 * m21 * m21 - m11 * m20 * m20 - m22 * m10 * m10;
  const float c1 = m00 * m11 - m10 * m10 + m00 * m22 - m20 * m20 + m11 * m22 - m21 * m21;
  const float c2 = m00 + m11 + m22;
  const float c2_3 = c2 * inv3;
  float a_3 = (c2 * c2_3 - c1) * inv3;
  a_3 = a_3 > 0.0f ? a_3 : 0.0f;
  const float half_b = 0.5f * (c0 + c2_3 * (2.0f * c2_3 * c2_3 - c1));
  float q = a_3 * a_3 * a_3 - half_b * half_b;
  q = q > 0.0f ? q : 0.0f;
  const float rho = sqrtf(a_3);
  const float theta = atan2f(sqrtf(q), half_b) * inv3;
  const float ct = cosf(theta), st = sinf(theta);
  w[0] = c2_3 - rho * (ct + sqrt3 * st);
  w[1] = c2_3 - rho * (ct - sqrt3 * st);
  w[2] = c2_3 + 2.0f * rho * ct;
  float E[3][3];  // E[k]: eigenvector k
  const float eps = 1.1920928955078125e-07f;  // FLT_EPSILON
  if (w[2] - w[0] <= eps) {
    for (int k = 0; k < 3; ++k)
      for (int r = 0; r < 3; ++r) E[k][r] = k == r ? 1.0f : 0.0f;
  } else {
    float d0 = w[2] - w[1];
    const float d1 = w[1] - w[0];
    int k = 0, l = 2;
    if (d0 > d1) {
      k = 2;
      l = 0;
      d0 = d1;
    }
    float T[3][3];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) T[r][c] = S[r][c];
    for (int i = 0; i < 3; ++i) T[i][i] -= w[k];
    kernel_vector(T, E[k], E[l]);
    if (d0 <= 2.0f * eps * d1) {
      // the other two coincide: the saved column, with the projection taken out as Eigen writes it
      const float pr = dot3f(E[k], E[l]);
      for (int r = 0; r < 3; ++r) E[l][r] -= pr * E[l][r];
      const float z = dot3f(E[l], E[l]);
      if (z > 0.0f) {
        const float d = sqrtf(z);
        for (int r = 0; r < 3; ++r) E[l][r] /= d;
      }
    } else {
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) T[r][c] = S[r][c];
      for (int i = 0; i < 3; ++i) T[i][i] -= w[l];
      float unused[3];
      kernel_vector(T, E[l], unused);
    }
    float m[3];
    cross3f(E[2], E[0], m);
    const float z = dot3f(m, m);
    const float d = z > 0.0f ? sqrtf(z) : 1.0f;
    for (int r = 0; r < 3; ++r) E[1][r] = z > 0.0f ? m[r] / d : m[r];
  }
  for (int k = 0; k < 3; ++k) w[k] = w[k] * scale + shift;
  for (int k = 0; k < 3; ++k)
    for (int r = 0; r < 3; ++r) V[r][k] = E[k][r];
}

// fillMetadata's sums for label_mode 1 over the points of one node, in the
// node's point order (xyz: 3 floats per point, stamp: per point, ts_start:
// the root's smallest stamp): mean, FP32 covariance and the stamp moments.
struct RefMoments {
  float mean[3];
  float cov[3][3];
  double dts_mean, dts_cov;
};
template <typename GetPoint>
inline void ref_moments(long n, const GetPoint& get, double ts_start, RefMoments& M) {
  LevelAdder<float, 9> acc;
  LevelAdder<double, 2> tacc;
  for (long i = 0; i < n; ++i) {
    float x[3];
    double st;
    get(i, x, st);
    const double dts = st - ts_start;
    const double t2[2] = {dts, dts * dts};
    tacc.add(t2);
    float packed[9];
    int k = 0;
    for (int r = 0; r < 3; ++r)
      for (int c = r; c < 3; ++c, ++k) packed[k] = x[r] * x[c];
    for (int r = 0; r < 3; ++r, ++k) packed[k] = x[r];
    acc.add(packed);
  }
  const float isize = 1.0f / (float)n;
  double ts[2];
  tacc.sum(ts);
  const double t1 = ts[0] * (double)isize, t2 = ts[1] * (double)isize;
  M.dts_mean = t1;
  M.dts_cov = t2 - t1 * t1;
  float packed[9];
  acc.sum(packed);
  for (int k = 0; k < 9; ++k) packed[k] *= isize;
  int k = 0;
  for (int r = 0; r < 3; ++r)
    for (int c = r; c < 3; ++c, ++k) M.cov[r][c] = M.cov[c][r] = packed[k];
  for (int r = 0; r < 3; ++r, ++k) M.mean[r] = packed[k];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) M.cov[r][c] -= M.mean[r] * M.mean[c];
}

// fillMetadata then the policies, label_mode 1.
inline void label_node_ref(const kdh_tree_gen_params& P, long n, const RefMoments& M, bool root,
                           const float inherited_prop[3], int inherited_has, Label& L) {
  L.dts_mean = M.dts_mean;
  L.dts_cov = M.dts_cov;
  for (int k = 0; k < 3; ++k) L.mean[k] = M.mean[k];
  bool is_bad = true;
  float evec[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  double evals[3] = {0, 0, 0};
  if (n >= 3) {
    float w[3], V[3][3];
    eig3_direct(M.cov, w, V);
    for (int k = 0; k < 3; ++k) {
      evals[k] = (double)w[k];
      for (int i = 0; i < 3; ++i) evec[k][i] = V[i][k];
    }
    if (dot3f(L.mean, evec[0]) > 0)
      for (int i = 0; i < 3; ++i) evec[0][i] = -evec[0][i];
    if (dot3f(L.mean, evec[2]) > 0)
      for (int i = 0; i < 3; ++i) evec[2][i] = -evec[2][i];
    is_bad = false;
  }
  label_policies(P, n, is_bad, evals, evec, root, inherited_prop, inherited_has, L);
}

// Node_::whichSide (tree/node_.h:54-57): left when direction . (p - mean) > 0.
KDH_HD bool goes_left(const float dir[3], const float mean[3], float x, float y, float z) {
  const float d[3] = {x - mean[0], y - mean[1], z - mean[2]};
  return dot3f(dir, d) > 0;
}

// IsGoodLeafPredicate_ (tree/tree_predicates_.h:8-29) with its defaults.
KDH_HD bool good_leaf(const float mean[3], const float dir[3]) {
  const float sq = (dir[0] * dir[0] + dir[1] * dir[1]) + dir[2] * dir[2];
  if (!(sq > 0)) return false;
  return fabsf(dot3f(mean, dir)) > 0.1f;
}

}  // namespace tg
}  // namespace kdh
