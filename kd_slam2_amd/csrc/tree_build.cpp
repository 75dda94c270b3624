// Host-side PCA KD-tree builder: the input preparation of the alignment path.
//
// Restates TreeGenerator_::makeTree (kd_slam2/src/lib/kd_slam/tree/tree_generator_impl_.h:9-19):
//   recursive PCA split through the node mean along the largest eigenvector
//     (Node_ constructor and splitRange, tree/node_impl_.h:13-70),
//   node labelling by FlatLeafPolicy_ over CountLeafPolicy_ over
//     LabelPolicyBase_ (flat_leaf_policy_.h:52-108, count_leaf_policy_.h:27-43,
//     label_policy_base_impl_.h:13-97),
//   prune with IsGoodLeafPredicate_ (tree_predicates_.h:8-29, tree_cpu_.h:173-207),
//   breadth-first renumbering (tree_cpu_impl_.h:48-74, 137-158) and the leaf list.
// Differences from the reference, none of which change what the aligner reads
// from a given tree:
//   * node mean, covariance and stamp moments are computed in FP64 with a
//     centred two-pass formula (the reference sums E[xx^T] in FP32 and
//     subtracts mu mu^T, which cancels badly far from the origin), then
//     rounded to FP32;
//   * the 3x3 symmetric eigenproblem is solved by cyclic Jacobi in FP64
//     (the reference calls Eigen's SelfAdjointEigenSolver::computeDirect);
//   * a split that leaves one side empty turns the node into a leaf (the
//     reference would recurse on the same range without end).
#include <math.h>
#include <string.h>

#include <algorithm>
#include <deque>
#include <future>
#include <limits>
#include <memory>
#include <vector>

#include "../../include/kdh.h"
#include "../../include/kdh_host.h"

namespace {

struct Pt {
  float x, y, z;
  double stamp;
};

struct Node {
  kdh_node3f n;
};

struct Meta {
  double mean[3] = {0, 0, 0};
  float meanf[3] = {0, 0, 0};
  double evals[3] = {0, 0, 0};  // ascending
  float evec[3][3] = {};        // evec[k] = k-th eigenvector (ascending eigenvalue)
  size_t n_points = 0;
  bool is_node = false, is_leaf = false, is_bad = true;
  int level = 0;
  double ts_start = 0;
  bool is_flat = false, has_propagated_normal = false;
  float propagated_normal[3] = {0, 0, 0};
};

static inline float dotf(const float a[3], const float b[3]) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

// Cyclic Jacobi eigen-decomposition of a symmetric 3x3 matrix (FP64).
// Returns eigenvalues ascending and unit eigenvectors as columns of V.
static void eig3(const double A_in[3][3], double w[3], double V[3][3]) {
  double A[3][3];
  memcpy(A, A_in, sizeof(A));
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) V[i][j] = (i == j) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 50; ++sweep) {
    double off = fabs(A[0][1]) + fabs(A[0][2]) + fabs(A[1][2]);
    double scale = fabs(A[0][0]) + fabs(A[1][1]) + fabs(A[2][2]);
    if (off <= 1e-300 || off <= 1e-18 * scale) break;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        if (fabs(A[p][q]) <= 1e-300) continue;
        double theta = (A[q][q] - A[p][p]) / (2.0 * A[p][q]);
        double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < 3; ++k) {  // A <- J^T A J
          double akp = A[k][p], akq = A[k][q];
          A[k][p] = c * akp - s * akq;
          A[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < 3; ++k) {
          double apk = A[p][k], aqk = A[q][k];
          A[p][k] = c * apk - s * aqk;
          A[q][k] = s * apk + c * aqk;
        }
        for (int k = 0; k < 3; ++k) {
          double vkp = V[k][p], vkq = V[k][q];
          V[k][p] = c * vkp - s * vkq;
          V[k][q] = s * vkp + c * vkq;
        }
      }
  }
  int idx[3] = {0, 1, 2};
  double d[3] = {A[0][0], A[1][1], A[2][2]};
  std::sort(idx, idx + 3, [&](int a, int b) { return d[a] < d[b]; });
  double Vs[3][3];
  for (int k = 0; k < 3; ++k) {
    w[k] = d[idx[k]];
    for (int i = 0; i < 3; ++i) Vs[i][k] = V[i][idx[k]];
  }
  memcpy(V, Vs, sizeof(Vs));
}

struct Builder {
  const kdh_tree_gen_params& P;
  Pt* pts;
  std::vector<Node> nodes;
  explicit Builder(const kdh_tree_gen_params& p, Pt* points) : P(p), pts(points) {}

  // LabelPolicyBase_::fillMetadata (label_policy_base_impl_.h:13-97)
  void fill_base(Meta& m, kdh_node3f& node, const Meta* parent) const {
    const size_t s = node.idx_start, e = node.idx_end;
    m.n_points = e - s;
    if (!m.n_points) {
      m.is_bad = true;
      return;
    }
    if (!parent) {
      double ts = std::numeric_limits<double>::max();
      for (size_t i = s; i < e; ++i) ts = std::min(ts, pts[i].stamp);
      m.ts_start = ts;
    } else {
      m.ts_start = parent->ts_start;
    }
    m.level = parent ? parent->level + 1 : 0;
    const double inv = 1.0 / (double)m.n_points;
    double t1 = 0, t2 = 0;
    for (size_t i = s; i < e; ++i) {
      double dts = pts[i].stamp - m.ts_start;
      t1 += dts;
      t2 += dts * dts;
    }
    t1 *= inv;
    t2 *= inv;
    node.dts_mean = t1;
    node.dts_covariance = t2 - t1 * t1;
    double mu[3] = {0, 0, 0};
    for (size_t i = s; i < e; ++i) {
      mu[0] += pts[i].x;
      mu[1] += pts[i].y;
      mu[2] += pts[i].z;
    }
    for (int k = 0; k < 3; ++k) {
      mu[k] *= inv;
      m.mean[k] = mu[k];
      m.meanf[k] = (float)mu[k];
      node.mean[k] = m.meanf[k];
    }
    if (m.n_points < 3) {
      node.direction[0] = node.direction[1] = node.direction[2] = 0.0f;
      m.is_bad = true;
      return;
    }
    double C[3][3] = {};
    for (size_t i = s; i < e; ++i) {
      double d[3] = {pts[i].x - mu[0], pts[i].y - mu[1], pts[i].z - mu[2]};
      for (int r = 0; r < 3; ++r)
        for (int c = r; c < 3; ++c) C[r][c] += d[r] * d[c];
    }
    for (int r = 0; r < 3; ++r)
      for (int c = r; c < 3; ++c) {
        C[r][c] *= inv;
        C[c][r] = C[r][c];
      }
    double V[3][3];
    eig3(C, m.evals, V);
    for (int k = 0; k < 3; ++k)
      for (int i = 0; i < 3; ++i) m.evec[k][i] = (float)V[i][k];
    // reorient the normal (smallest) and the direction (largest) against the mean
    if (dotf(m.meanf, m.evec[0]) > 0)
      for (int i = 0; i < 3; ++i) m.evec[0][i] = -m.evec[0][i];
    if (dotf(m.meanf, m.evec[2]) > 0)
      for (int i = 0; i < 3; ++i) m.evec[2][i] = -m.evec[2][i];
    m.is_bad = false;
  }

  static void set_dir(kdh_node3f& node, const float d[3]) {
    node.direction[0] = d[0];
    node.direction[1] = d[1];
    node.direction[2] = d[2];
  }

  // FlatLeafPolicy_::propagateNormal (flat_leaf_policy_.h:38-50)
  bool propagate(kdh_node3f& node, Meta& m) const {
    m.is_leaf = true;
    if (m.has_propagated_normal) {
      set_dir(node, m.propagated_normal);
      if (dotf(node.mean, node.direction) > 0)
        for (int i = 0; i < 3; ++i) node.direction[i] = -node.direction[i];
      return true;
    }
    return false;
  }

  // CountLeafPolicy_ + FlatLeafPolicy_::fillMetadata
  void fill(Meta& m, kdh_node3f& node, const Meta* parent) const {
    fill_base(m, node, parent);
    // CountLeafPolicy_ (count_leaf_policy_.h:27-43)
    if (m.is_bad) {
      m.is_leaf = true;
    } else if ((double)m.n_points > (double)P.leaf_max_points) {
      m.is_node = true;
      set_dir(node, m.evec[2]);
    } else {
      m.is_leaf = true;
      set_dir(node, m.evec[0]);
    }
    // FlatLeafPolicy_ (flat_leaf_policy_.h:52-108)
    if (parent && P.propagate_normal) {
      memcpy(m.propagated_normal, parent->propagated_normal, sizeof(m.propagated_normal));
      m.has_propagated_normal = parent->has_propagated_normal;
    } else {
      m.propagated_normal[0] = m.propagated_normal[1] = m.propagated_normal[2] = 0.0f;
      m.has_propagated_normal = false;
    }
    if (m.is_bad || (double)m.n_points < (double)P.leaf_min_points) {
      m.is_bad = !propagate(node, m);
      return;
    }
    const float lmax = (float)m.evals[2], lmin = (float)m.evals[0];
    const float ratio = lmin / lmax;
    m.is_flat = ratio < P.leaf_normal_eigenratio && fabsf(dotf(node.mean, m.evec[0])) > P.leaf_min_normal_incidence;
    if (m.is_flat) {
      m.has_propagated_normal = true;
      memcpy(m.propagated_normal, m.evec[0], sizeof(m.propagated_normal));
    }
    if (m.is_node) return;
    float volume = 1.0f;
    for (int i = 1; i < 3; ++i) volume *= (float)m.evals[i];
    const float density = (float)m.n_points / volume;
    if (density < P.leaf_density) {
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
    m.is_bad = !propagate(node, m);
  }

  // splitRange (node_impl_.h:13-26): left = direction . (p - mean) > 0
  size_t split(const kdh_node3f& node) {
    long s = (long)node.idx_start, e = (long)node.idx_end - 1;
    while (s <= e) {
      const Pt& p = pts[s];
      float d[3] = {p.x - node.mean[0], p.y - node.mean[1], p.z - node.mean[2]};
      bool left = dotf(node.direction, d) > 0;
      if (left) {
        ++s;
      } else {
        std::swap(pts[s], pts[e]);
        --e;
      }
    }
    return (size_t)s;
  }

  int build(size_t start, size_t end, const Meta* parent) {
    const int me = (int)nodes.size();
    nodes.push_back(Node{});
    kdh_node3f node;
    memset(&node, 0, sizeof(node));
    node.idx_start = start;
    node.idx_end = end;
    node.idx_left = -1;
    node.idx_right = -1;
    Meta m;
    fill(m, node, parent);
    if (!m.is_node) {
      nodes[me].n = node;
      return me;
    }
    const size_t mid = split(node);
    if (mid == start || mid == end) {
      // degenerate split: keep the node as a leaf with its normal
      set_dir(node, m.evec[0]);
      nodes[me].n = node;
      return me;
    }
    nodes[me].n = node;
    int l, r;
    if (P.parallel_level > 0 && m.level < P.parallel_level) {
      Builder bl(P, pts), br(P, pts);
      auto fl = std::async(std::launch::async, [&] { return bl.build(start, mid, &m); });
      br.build(mid, end, &m);
      fl.get();
      l = append(bl);
      r = append(br);
    } else {
      l = build(start, mid, &m);
      r = build(mid, end, &m);
    }
    nodes[me].n.idx_left = l;
    nodes[me].n.idx_right = r;
    return me;
  }

  int append(const Builder& sub) {
    const int off = (int)nodes.size();
    for (const Node& s : sub.nodes) {
      Node n = s;
      if (n.n.idx_left >= 0) n.n.idx_left += off;
      if (n.n.idx_right >= 0) n.n.idx_right += off;
      nodes.push_back(n);
    }
    return off;
  }
};

static bool is_leaf(const kdh_node3f& n) { return n.idx_left < 0 && n.idx_right < 0; }

// IsGoodLeafPredicate_ (tree_predicates_.h:8-29) with its defaults
// (min_normal_incidence 0.1; its min_points is unused by the reference).
static bool good_leaf(const kdh_node3f& n) {
  if (!is_leaf(n)) return false;
  const float sq = (n.direction[0] * n.direction[0] + n.direction[1] * n.direction[1]) + n.direction[2] * n.direction[2];
  if (!(sq > 0)) return false;
  const float inc = dotf(n.mean, n.direction);
  if (fabsf(inc) <= 0.1f) return false;
  return true;
}

// TreeCPU_::_prune (tree_cpu_.h:209-232)
static int prune_rec(std::vector<Node>& nodes, int idx) {
  kdh_node3f& n = nodes[idx].n;
  if (is_leaf(n)) return good_leaf(n) ? idx : -1;
  if (n.idx_left >= 0) n.idx_left = prune_rec(nodes, n.idx_left);
  if (n.idx_right >= 0) n.idx_right = prune_rec(nodes, n.idx_right);
  if (is_leaf(n)) {
    n.direction[0] = n.direction[1] = n.direction[2] = 0.0f;
    return -1;
  }
  return idx;
}

}  // namespace

struct kdh_host_tree {
  std::vector<kdh_node3f> nodes;
  std::vector<uint64_t> leaves;
  std::vector<float> points;   // permuted points (x, y, z)
  std::vector<double> stamps;  // permuted stamps
  double ts_start = 0;
};

extern "C" {

void kdh_tree_gen_params_default(kdh_tree_gen_params* p) {
  if (!p) return;
  p->propagate_normal = 0;
  p->tree_min_good_leaves = 10;
  p->leaf_min_points = 5;
  p->leaf_max_points = 30;
  p->leaf_density = 100.0f;
  p->leaf_min_normal_incidence = 0.1f;
  p->leaf_eigenvalue_threshold = 1e-3f;
  p->leaf_normal_eigenratio = 0.1f;
  p->parallel_level = 0;
}

int kdh_host_tree_build(const float* xyz, const double* stamps, int64_t n_points, const kdh_tree_gen_params* params,
                        kdh_host_tree** out) {
  if (!out || !params || n_points < 0 || (n_points > 0 && !xyz)) return KDH_ERR_INVALID;
  *out = nullptr;
  const kdh_tree_gen_params& P = *params;
  if (n_points < (int64_t)P.leaf_min_points) return KDH_OK;  // makeTree returns nullptr
  std::vector<Pt> pts((size_t)n_points);
  for (int64_t i = 0; i < n_points; ++i) {
    pts[i].x = xyz[3 * i];
    pts[i].y = xyz[3 * i + 1];
    pts[i].z = xyz[3 * i + 2];
    pts[i].stamp = stamps ? stamps[i] : 0.0;
  }
  Builder b(P, pts.data());
  b.nodes.reserve((size_t)n_points / 4 + 16);
  b.build(0, (size_t)n_points, nullptr);
  double ts_start = std::numeric_limits<double>::max();
  for (const Pt& p : pts) ts_start = std::min(ts_start, p.stamp);

  std::vector<Node>& nodes = b.nodes;
  int root = prune_rec(nodes, 0);
  auto* t = new kdh_host_tree();
  t->ts_start = ts_start;
  if (root >= 0) {
    // breadthFirstPermutation + applyPermutation
    std::vector<int> perm;
    perm.reserve(nodes.size());
    std::deque<int> frontier;
    frontier.push_back(0);
    while (!frontier.empty()) {
      int cur = frontier.front();
      frontier.pop_front();
      perm.push_back(cur);
      const kdh_node3f& n = nodes[cur].n;
      if (n.idx_left >= 0) frontier.push_back(n.idx_left);
      if (n.idx_right >= 0) frontier.push_back(n.idx_right);
    }
    std::vector<int> inv(nodes.size(), -1);
    for (size_t i = 0; i < perm.size(); ++i) inv[perm[i]] = (int)i;
    t->nodes.resize(perm.size());
    for (size_t i = 0; i < perm.size(); ++i) {
      kdh_node3f n = nodes[perm[i]].n;
      if (n.idx_left >= 0) n.idx_left = inv[n.idx_left];
      if (n.idx_right >= 0) n.idx_right = inv[n.idx_right];
      t->nodes[i] = n;
    }
    // recomputeLeaves
    for (size_t i = 0; i < t->nodes.size(); ++i)
      if (is_leaf(t->nodes[i])) t->leaves.push_back(i);
  }
  if ((int64_t)t->leaves.size() < (int64_t)P.tree_min_good_leaves) {
    delete t;
    return KDH_OK;  // makeTree returns nullptr
  }
  t->points.resize((size_t)n_points * 3);
  t->stamps.resize((size_t)n_points);
  for (int64_t i = 0; i < n_points; ++i) {
    t->points[3 * i] = pts[i].x;
    t->points[3 * i + 1] = pts[i].y;
    t->points[3 * i + 2] = pts[i].z;
    t->stamps[i] = pts[i].stamp;
  }
  *out = t;
  return KDH_OK;
}

int kdh_host_tree_info(const kdh_host_tree* t, int64_t* n_nodes, int64_t* n_leaves, double* ts_start) {
  if (!t) return KDH_ERR_INVALID;
  if (n_nodes) *n_nodes = (int64_t)t->nodes.size();
  if (n_leaves) *n_leaves = (int64_t)t->leaves.size();
  if (ts_start) *ts_start = t->ts_start;
  return KDH_OK;
}

const kdh_node3f* kdh_host_tree_nodes(const kdh_host_tree* t) { return t ? t->nodes.data() : nullptr; }
const uint64_t* kdh_host_tree_leaves(const kdh_host_tree* t) { return t ? t->leaves.data() : nullptr; }
const float* kdh_host_tree_points(const kdh_host_tree* t) { return t ? t->points.data() : nullptr; }

int kdh_host_tree_free(kdh_host_tree* t) {
  if (!t) return KDH_ERR_INVALID;
  delete t;
  return KDH_OK;
}

}  // extern "C"
