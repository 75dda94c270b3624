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

// LabelPolicyBase_::fillMetadata after the sums, then CountLeafPolicy_ and
// FlatLeafPolicy_.  n: points; s1 = (sum x, sum y, sum z, sum dts, sum dts^2)
// in the canonical order; c2 = centred sums (xx, xy, xz, yy, yz, zz) about the
// FP64 mean s1[0..2] / n in the canonical order.  inherited prop / has_prop:
// the parent's (ignored for the root or without propagate_normal).
KDH_HD void label_node(const kdh_tree_gen_params& P, long n, const double s1[5], const double c2[6], bool root,
                       const float inherited_prop[3], int inherited_has, Label& L) {
  const double inv = 1.0 / (double)n;
  const double t1 = s1[3] * inv, t2 = s1[4] * inv;
  L.dts_mean = t1;
  L.dts_cov = t2 - t1 * t1;
  for (int k = 0; k < 3; ++k) L.mean[k] = (float)(s1[k] * inv);
  for (int k = 0; k < 3; ++k) L.dir[k] = L.evec0[k] = 0.0f;
  bool is_bad = true, is_leaf = false, is_node = false;
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
    for (int i = 0; i < 3; ++i) L.evec0[i] = evec[0][i];
  }
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
