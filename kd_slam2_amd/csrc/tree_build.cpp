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
// The moments, the eigen-solver, the labelling and the split rule are shared
// with the device builder (kdh_treegen.h, which lists where this build fixes
// what the reference leaves open); both give the same tree bit for bit.
// label_mode 1 instead evaluates them with the reference's own arithmetic
// (FP32 StableAdder_ moments, computeDirect, the two-ended splitRange); the
// device builder has no such mode.
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
#include "kdh_treegen.h"

namespace {

using kdh::tg::Label;

struct Pt {
  float x, y, z;
  double stamp;
};

struct Builder {
  const kdh_tree_gen_params& P;
  Pt* pts;
  double ts_start;
  std::vector<kdh_node3f> nodes;
  std::vector<Pt> tmp;
  std::vector<double> acc;
  Builder(const kdh_tree_gen_params& p, Pt* points, double ts0) : P(p), pts(points), ts_start(ts0) {}

  // The node's moments in the canonical order: s1 = sums of (x, y, z, dts,
  // dts^2); c2 = sums of the centred products (xx, xy, xz, yy, yz, zz).
  void moments(size_t s, size_t e, double s1[5], double c2[6]) {
    const size_t n = e - s;
    acc.resize(n * 6);
    double(*v5)[5] = reinterpret_cast<double(*)[5]>(acc.data());
    for (size_t i = s; i < e; ++i) {
      const double dts = pts[i].stamp - ts_start;
      double* a = v5[i - s];
      a[0] = (double)pts[i].x;
      a[1] = (double)pts[i].y;
      a[2] = (double)pts[i].z;
      a[3] = dts;
      a[4] = dts * dts;
    }
    kdh::tg::reduce_canonical<5>(v5, (long)n);
    for (int k = 0; k < 5; ++k) s1[k] = v5[0][k];
    const double inv = 1.0 / (double)n;
    const double mu[3] = {s1[0] * inv, s1[1] * inv, s1[2] * inv};
    double(*v6)[6] = reinterpret_cast<double(*)[6]>(acc.data());
    for (size_t i = s; i < e; ++i) {
      const double d0 = (double)pts[i].x - mu[0], d1 = (double)pts[i].y - mu[1], d2 = (double)pts[i].z - mu[2];
      double* a = v6[i - s];
      a[0] = d0 * d0;
      a[1] = d0 * d1;
      a[2] = d0 * d2;
      a[3] = d1 * d1;
      a[4] = d1 * d2;
      a[5] = d2 * d2;
    }
    kdh::tg::reduce_canonical<6>(v6, (long)n);
    for (int k = 0; k < 6; ++k) c2[k] = v6[0][k];
  }

  // splitRange as the reference writes it (tree/node_impl_.h:12-26): a point
  // on the right side is swapped with the last unexamined one.  Signed
  // indices: an all-right range starting at point 0 ends the scan at s = 0
  // (the reference's unsigned end index would wrap there).
  size_t split_ref(const kdh_node3f& node) {
    long s = (long)node.idx_start, e = (long)node.idx_end - 1;
    while (s <= e) {
      if (kdh::tg::goes_left(node.direction, node.mean, pts[s].x, pts[s].y, pts[s].z)) {
        ++s;
      } else {
        std::swap(pts[s], pts[e]);
        --e;
      }
    }
    return (size_t)s;
  }

  // splitRange with the order of each side kept
  size_t split(const kdh_node3f& node) {
    const size_t s = node.idx_start, e = node.idx_end;
    tmp.clear();
    size_t w = s;
    for (size_t i = s; i < e; ++i) {
      if (kdh::tg::goes_left(node.direction, node.mean, pts[i].x, pts[i].y, pts[i].z))
        pts[w++] = pts[i];
      else
        tmp.push_back(pts[i]);
    }
    std::copy(tmp.begin(), tmp.end(), pts + w);
    return w;
  }

  int build(size_t start, size_t end, const Label* parent, int level) {
    const int me = (int)nodes.size();
    nodes.push_back(kdh_node3f{});
    kdh_node3f node;
    memset(&node, 0, sizeof(node));
    node.idx_start = start;
    node.idx_end = end;
    node.idx_left = -1;
    node.idx_right = -1;
    static const float zero3[3] = {0.0f, 0.0f, 0.0f};
    Label L;
    if (P.label_mode == 1) {
      kdh::tg::RefMoments M;
      const Pt* base = pts + start;
      kdh::tg::ref_moments(
          (long)(end - start),
          [base](long i, float x[3], double& st) {
            x[0] = base[i].x;
            x[1] = base[i].y;
            x[2] = base[i].z;
            st = base[i].stamp;
          },
          ts_start, M);
      kdh::tg::label_node_ref(P, (long)(end - start), M, parent == nullptr, parent ? parent->prop : zero3,
                              parent ? parent->has_prop : 0, L);
    } else {
      double s1[5], c2[6];
      moments(start, end, s1, c2);
      kdh::tg::label_node(P, (long)(end - start), s1, c2, parent == nullptr, parent ? parent->prop : zero3,
                          parent ? parent->has_prop : 0, L);
    }
    node.dts_mean = L.dts_mean;
    node.dts_covariance = L.dts_cov;
    memcpy(node.mean, L.mean, sizeof(node.mean));
    memcpy(node.direction, L.dir, sizeof(node.direction));
    if (!L.is_node) {
      nodes[me] = node;
      return me;
    }
    const size_t mid = P.label_mode == 1 ? split_ref(node) : split(node);
    if (mid == start || mid == end) {
      // one-sided split: keep the node as a leaf with its normal
      memcpy(node.direction, L.evec0, sizeof(node.direction));
      nodes[me] = node;
      return me;
    }
    nodes[me] = node;
    int l, r;
    if (P.parallel_level > 0 && level < P.parallel_level) {
      Builder bl(P, pts, ts_start), br(P, pts, ts_start);
      auto fl = std::async(std::launch::async, [&] { return bl.build(start, mid, &L, level + 1); });
      br.build(mid, end, &L, level + 1);
      fl.get();
      l = append(bl);
      r = append(br);
    } else {
      l = build(start, mid, &L, level + 1);
      r = build(mid, end, &L, level + 1);
    }
    nodes[me].idx_left = l;
    nodes[me].idx_right = r;
    return me;
  }

  int append(const Builder& sub) {
    const int off = (int)nodes.size();
    for (const kdh_node3f& s : sub.nodes) {
      kdh_node3f n = s;
      if (n.idx_left >= 0) n.idx_left += off;
      if (n.idx_right >= 0) n.idx_right += off;
      nodes.push_back(n);
    }
    return off;
  }
};

static bool is_leaf(const kdh_node3f& n) { return n.idx_left < 0 && n.idx_right < 0; }

// IsGoodLeafPredicate_ (tree_predicates_.h:8-29) with its defaults
// (min_normal_incidence 0.1; its min_points is unused by the reference).
static bool good_leaf(const kdh_node3f& n) { return is_leaf(n) && kdh::tg::good_leaf(n.mean, n.direction); }

// TreeCPU_::_prune (tree_cpu_.h:209-232)
static int prune_rec(std::vector<kdh_node3f>& nodes, int idx) {
  kdh_node3f& n = nodes[idx];
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
  p->label_mode = 0;
}

int kdh_host_tree_build(const float* xyz, const double* stamps, int64_t n_points, const kdh_tree_gen_params* params,
                        kdh_host_tree** out) {
  if (!out || !params || n_points < 0 || (n_points > 0 && !xyz)) return KDH_ERR_INVALID;
  if (params->label_mode != 0 && params->label_mode != 1) return KDH_ERR_INVALID;
  *out = nullptr;
  const kdh_tree_gen_params& P = *params;
  // makeTree returns nullptr below leaf_min_points; an empty cloud gives no tree
  if (n_points < (int64_t)P.leaf_min_points || n_points == 0) return KDH_OK;
  std::vector<Pt> pts((size_t)n_points);
  double ts_start = std::numeric_limits<double>::max();
  for (int64_t i = 0; i < n_points; ++i) {
    pts[i].x = xyz[3 * i];
    pts[i].y = xyz[3 * i + 1];
    pts[i].z = xyz[3 * i + 2];
    pts[i].stamp = stamps ? stamps[i] : 0.0;
    ts_start = std::min(ts_start, pts[i].stamp);
  }
  Builder b(P, pts.data(), ts_start);
  b.nodes.reserve((size_t)n_points / 4 + 16);
  b.build(0, (size_t)n_points, nullptr, 0);

  std::vector<kdh_node3f>& nodes = b.nodes;
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
      const kdh_node3f& n = nodes[cur];
      if (n.idx_left >= 0) frontier.push_back(n.idx_left);
      if (n.idx_right >= 0) frontier.push_back(n.idx_right);
    }
    std::vector<int> inv(nodes.size(), -1);
    for (size_t i = 0; i < perm.size(); ++i) inv[perm[i]] = (int)i;
    t->nodes.resize(perm.size());
    for (size_t i = 0; i < perm.size(); ++i) {
      kdh_node3f n = nodes[perm[i]];
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
