// PCA KD-tree generation on the device (kdh_tree_build): TreeGenerator_::
// makeTree (tree/tree_generator_impl_.h:9-19) followed by the device upload
// (Tree_CUDAfromCPU_, tree/tree_cuda_impl_.h:198-222), without a host round
// trip.  The tree is bit-identical to the host builder's (tree_build.cpp):
// both evaluate the shared labelling, eigen-solver and split rule of
// kdh_treegen.h, sum node moments in its canonical order and keep the point
// order of each side of a split.
//
// The build runs level by level.  Level L holds a contiguous range of nodes;
// their points are cut into chunks of 64 (a chunk never straddles two
// nodes), and every pass over the points is one wave per chunk, one lane per
// point, loads coalesced:
//   mom1    wave per chunk: pairwise sums of (x, y, z, dts, dts^2)
//   red1    wave per node:  the chunk sums in the canonical pairwise order
//   mom2    wave per chunk: the products centred on the FP64 mean
//   red2    wave per node:  their reduction, the eigenproblem, the labels
//   count   wave per chunk: points going left (ballot)
//   level   one workgroup:  prefix sums -> split sizes, children, next chunks
//   scatter wave per chunk: stable partition into the other point buffer
//                           (ballot ranks)
// The node and chunk counts stay on the device; the host queues levels
// ahead and reads a status word every few levels.  Then one workgroup
// prunes bottom-up (IsGoodLeafPredicate_) and numbers the kept nodes in
// breadth-first order, and a last pass writes the device tree layout of
// kdh_tree_upload plus the full Node_ records, the leaf list and the
// permuted points.
#include <hip/hip_runtime.h>

#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/kdh.h"
#include "../../include/kdh_host.h"
#include "kdh_capi_internal.h"
#include "kdh_treegen.h"

#pragma clang fp contract(off)

using namespace kdh;

#define KDH_HIP_TRY(ctx, expr)                                 \
  do {                                                         \
    hipError_t _e = (expr);                                    \
    if (_e != hipSuccess) return kdh_hip_fail(ctx, _e, #expr); \
  } while (0)

namespace {

constexpr int kC = tg::kChunk;
constexpr int kLevelThreads = 1024;
constexpr int kMaxLevels = 4096;
constexpr int kBatchLevels = 4;  // levels queued between status checks

// One node of the level-ordered build (before pruning).
struct BNode {
  int start, end;     // point range
  int left, right;    // children (-1)
  int level;
  int is_node;        // the labelling asks for a split
  int split;          // split into two non-empty sides
  int has_prop;       // FlatLeafPolicy_ propagated normal (inherited, then own)
  int n_left;
  int keep;           // survives the prune
  int new_index;      // breadth-first index after the prune
  int leaf_index;     // position in the leaf list
  float mean[3], dir[3], evec0[3], prop[3];
  double dts_mean, dts_cov;
  double s1[5];       // (x, y, z, dts, dts^2) sums
};

struct LevelCtl {
  int lb, le;         // nodes [lb, le)
  int n_chunks;
  int active;         // 1 while the level has nodes
};

struct Status {
  int last_level;     // first level with no nodes (-1 while building)
  int overflow;
  int n_out, n_leaves_out;
  int root_kept;
  int n_built;        // nodes of the unpruned tree
  double ts_start;
};

struct Ws {
  int64_t cap = 0;    // points
  float* px[2] = {nullptr, nullptr};
  float* py[2] = {nullptr, nullptr};
  float* pz[2] = {nullptr, nullptr};
  double* pt[2] = {nullptr, nullptr};
  BNode* nodes = nullptr;   // 2 * cap
  double* csum = nullptr;   // cap x 6
  int* cleft = nullptr;     // cap
  int* coff = nullptr;      // cap
  int* cbeg[2] = {nullptr, nullptr};  // cap + 1 each
  int* level_begin = nullptr;         // kMaxLevels + 2
  LevelCtl* ctl = nullptr;            // [2]
  Status* status = nullptr;
  float* stage_xyz = nullptr;         // 3 * cap (host inputs)
  double* stage_t = nullptr;          // cap
  Status* h_status = nullptr;         // pinned
  void* mem = nullptr;
};

__device__ __forceinline__ int find_node(const int* cbeg, int w, int c) {
  // last j in [0, w) with cbeg[j] <= c
  int lo = 0, hi = w - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (cbeg[mid] <= c)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Exclusive prefix sum over the 1024 threads of the block; *total gets the
// block sum.  s: 16 ints of LDS.
__device__ int block_exscan(int v, int* s, int* total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) s[wid] = x;
  __syncthreads();
  if (wid == 0) {
    int t = lane < 16 ? s[lane] : 0;
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) {
      const int y = __shfl_up(t, o, 64);
      if (lane >= o) t += y;
    }
    if (lane < 16) s[lane] = t;
  }
  __syncthreads();
  const int off = wid ? s[wid - 1] : 0;
  *total = s[15];
  __syncthreads();
  return off + x - v;
}

// ---- set-up ------------------------------------------------------------------

__global__ void k_stage(const float* __restrict__ xyz, const double* __restrict__ stamps, int n, float* px,
                        float* py, float* pz, double* pt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  px[i] = xyz[3 * i];
  py[i] = xyz[3 * i + 1];
  pz[i] = xyz[3 * i + 2];
  pt[i] = stamps ? stamps[i] : 0.0;
}

__global__ __launch_bounds__(kLevelThreads) void k_init(const double* __restrict__ pt, int n, BNode* nodes,
                                                        LevelCtl* ctl, int* cbeg0, int* level_begin, Status* st) {
  __shared__ double s_min[kLevelThreads];
  double m = 1.7976931348623157e308;
  for (int i = threadIdx.x; i < n; i += kLevelThreads) m = fmin(m, pt[i]);
  s_min[threadIdx.x] = m;
  __syncthreads();
  for (int s = kLevelThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) s_min[threadIdx.x] = fmin(s_min[threadIdx.x], s_min[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    BNode r;
    memset(&r, 0, sizeof(r));
    r.start = 0;
    r.end = n;
    r.left = r.right = -1;
    nodes[0] = r;
    const int nc = (n + kC - 1) / kC;
    ctl[0] = LevelCtl{0, 1, nc, 1};
    ctl[1] = LevelCtl{1, 1, 0, 0};
    cbeg0[0] = 0;
    cbeg0[1] = nc;
    level_begin[0] = 0;
    level_begin[1] = 1;
    Status s;
    memset(&s, 0, sizeof(s));
    s.last_level = -1;
    s.ts_start = s_min[0];
    *st = s;
  }
}

// ---- per level ---------------------------------------------------------------

struct LevelArgs {
  BNode* nodes;
  const LevelCtl* ctl;  // this level's
  const int* cbeg;      // this level's chunk starts (w + 1)
  const float *px, *py, *pz;
  const double* pt;
  double* csum;
  const Status* st;
};

// One group of the canonical order in a wave: lane i holds value i (i < cnt);
// lane 0 ends with the group sum.
template <int K>
__device__ __forceinline__ void wave_pairwise(double (&v)[K], int cnt) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int s = 1; s < tg::kGroup; s <<= 1) {
    const bool take = ((lane & (2 * s - 1)) == 0) && (lane + s < cnt);
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const double o = __shfl_down(v[k], s, 64);
      if (take) v[k] += o;
    }
  }
}

// The chunk of this wave: its node and point range.  Wave-uniform.
struct ChunkRef {
  int j;       // node within the level
  int p0;      // first point
  int cnt;     // points in the chunk (1..64)
};

__device__ __forceinline__ ChunkRef chunk_ref(const LevelArgs& a, const LevelCtl& L, int c) {
  ChunkRef r;
  r.j = find_node(a.cbeg, L.le - L.lb, c);
  const BNode& nd = a.nodes[L.lb + r.j];
  r.p0 = nd.start + (c - a.cbeg[r.j]) * kC;
  r.cnt = min(kC, nd.end - r.p0);
  return r;
}

__global__ void k_mom1(LevelArgs a) {
  const LevelCtl L = *a.ctl;
  const int c = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = threadIdx.x & 63;
  if (!L.active || c >= L.n_chunks) return;
  const ChunkRef r = chunk_ref(a, L, c);
  const int i = r.p0 + min(lane, r.cnt - 1);
  const double d = a.pt[i] - a.st->ts_start;
  double v[5] = {(double)a.px[i], (double)a.py[i], (double)a.pz[i], d, d * d};
  wave_pairwise<5>(v, r.cnt);
  if (lane == 0) {
    double* o = a.csum + (size_t)c * 6;
#pragma unroll
    for (int k = 0; k < 5; ++k) o[k] = v[k];
  }
}

__global__ void k_mom2(LevelArgs a) {
  const LevelCtl L = *a.ctl;
  const int c = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = threadIdx.x & 63;
  if (!L.active || c >= L.n_chunks) return;
  const ChunkRef r = chunk_ref(a, L, c);
  const BNode& nd = a.nodes[L.lb + r.j];
  const double inv = 1.0 / (double)(nd.end - nd.start);
  const double mu0 = nd.s1[0] * inv, mu1 = nd.s1[1] * inv, mu2 = nd.s1[2] * inv;
  const int i = r.p0 + min(lane, r.cnt - 1);
  const double d0 = (double)a.px[i] - mu0, d1 = (double)a.py[i] - mu1, d2 = (double)a.pz[i] - mu2;
  double v[6] = {d0 * d0, d0 * d1, d0 * d2, d1 * d1, d1 * d2, d2 * d2};
  wave_pairwise<6>(v, r.cnt);
  if (lane == 0) {
    double* o = a.csum + (size_t)c * 6;
#pragma unroll
    for (int k = 0; k < 6; ++k) o[k] = v[k];
  }
}

// The canonical reduction of the m chunk sums at csum[base..] by one wave;
// the result is returned in every lane.
template <int K>
__device__ void wave_reduce_chunks(double* csum, int base, int m, double out[K]) {
  const int lane = threadIdx.x & 63;
  while (m > 1) {
    const int groups = (m + tg::kGroup - 1) / tg::kGroup;
    for (int g = 0; g < groups; ++g) {
      const int b = g * tg::kGroup;
      const int cnt = min(m - b, tg::kGroup);
      double v[K];
      const double* src = csum + (size_t)(base + b + min(lane, cnt - 1)) * 6;
#pragma unroll
      for (int k = 0; k < K; ++k) v[k] = src[k];
      wave_pairwise<K>(v, cnt);
      // every lane has its value (the shuffles consumed them) before lane 0
      // overwrites slot g; the fences make the store visible to the next
      // round's loads by the other lanes
      if (lane == 0) {
        double* dst = csum + (size_t)(base + g) * 6;
#pragma unroll
        for (int k = 0; k < K; ++k) dst[k] = v[k];
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
    m = groups;
  }
  const double* r = csum + (size_t)base * 6;
#pragma unroll
  for (int k = 0; k < K; ++k) out[k] = r[k];
}

__global__ void k_red1(LevelArgs a) {
  const LevelCtl L = *a.ctl;
  if (!L.active) return;
  const int w = L.le - L.lb;
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int n_waves = (gridDim.x * blockDim.x) >> 6;
  for (int j = wave; j < w; j += n_waves) {
    const int base = a.cbeg[j], m = a.cbeg[j + 1] - base;
    double s[5];
    wave_reduce_chunks<5>(a.csum, base, m, s);
    if ((threadIdx.x & 63) == 0) {
      BNode& nd = a.nodes[L.lb + j];
#pragma unroll
      for (int k = 0; k < 5; ++k) nd.s1[k] = s[k];
    }
  }
}

__global__ void k_red2(LevelArgs a, kdh_tree_gen_params P) {
  const LevelCtl L = *a.ctl;
  if (!L.active) return;
  const int w = L.le - L.lb;
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int n_waves = (gridDim.x * blockDim.x) >> 6;
  for (int j = wave; j < w; j += n_waves) {
    const int base = a.cbeg[j], m = a.cbeg[j + 1] - base;
    double c2[6];
    wave_reduce_chunks<6>(a.csum, base, m, c2);
    if ((threadIdx.x & 63) == 0) {
      BNode& nd = a.nodes[L.lb + j];
      tg::Label lab;
      tg::label_node(P, (long)(nd.end - nd.start), nd.s1, c2, nd.level == 0, nd.prop, nd.has_prop, lab);
      for (int k = 0; k < 3; ++k) {
        nd.mean[k] = lab.mean[k];
        nd.dir[k] = lab.dir[k];
        nd.evec0[k] = lab.evec0[k];
        nd.prop[k] = lab.prop[k];
      }
      nd.has_prop = lab.has_prop;
      nd.is_node = lab.is_node;
      nd.dts_mean = lab.dts_mean;
      nd.dts_cov = lab.dts_cov;
    }
  }
}

__global__ void k_count(LevelArgs a, int* cleft) {
  const LevelCtl L = *a.ctl;
  const int c = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = threadIdx.x & 63;
  if (!L.active || c >= L.n_chunks) return;
  const ChunkRef r = chunk_ref(a, L, c);
  const BNode& nd = a.nodes[L.lb + r.j];
  bool left = false;
  if (nd.is_node && lane < r.cnt) {
    const int i = r.p0 + lane;
    left = tg::goes_left(nd.dir, nd.mean, a.px[i], a.py[i], a.pz[i]);
  }
  const unsigned long long m = __ballot(left);
  if (lane == 0) cleft[c] = __popcll(m);
}

// One workgroup: split sizes from the chunk counts, the children of the
// split nodes in breadth-first order, and the next level's chunk starts.
__global__ __launch_bounds__(kLevelThreads) void k_level(BNode* nodes, const LevelCtl* ctl, LevelCtl* ctl_next,
                                                         const int* cbeg, int* cbeg_next, const int* cleft,
                                                         int* coff, int* level_begin, int level, int cap_nodes,
                                                         Status* st) {
  __shared__ int s_scan[16];
  const LevelCtl L = *ctl;
  const int tid = threadIdx.x;
  if (!L.active) {
    if (tid == 0) *ctl_next = LevelCtl{L.le, L.le, 0, 0};
    return;
  }
  const int C = L.n_chunks, w = L.le - L.lb;
  int carry = 0;
  for (int b = 0; b < C; b += kLevelThreads) {
    const int c = b + tid;
    int tot;
    const int ex = block_exscan(c < C ? cleft[c] : 0, s_scan, &tot);
    if (c < C) coff[c] = carry + ex;
    carry += tot;
  }
  const int all_left = carry;
  __syncthreads();
  int carry_split = 0, carry_chunks = 0;
  for (int b = 0; b < w; b += kLevelThreads) {
    const int j = b + tid;
    int split = 0, nl = 0, cl = 0, cr = 0;
    if (j < w) {
      BNode& nd = nodes[L.lb + j];
      const int cb = cbeg[j], ce = cbeg[j + 1];
      nl = (ce < C ? coff[ce] : all_left) - coff[cb];
      const int n = nd.end - nd.start;
      split = (nd.is_node && nl > 0 && nl < n) ? 1 : 0;
      if (nd.is_node && !split)  // one-sided split: a leaf with its normal
        for (int k = 0; k < 3; ++k) nd.dir[k] = nd.evec0[k];
      nd.split = split;
      nd.n_left = nl;
      if (split) {
        cl = (nl + kC - 1) / kC;
        cr = (n - nl + kC - 1) / kC;
      }
    }
    int tot_split, tot_chunks;
    const int ks = block_exscan(split, s_scan, &tot_split);
    const int kc = block_exscan(cl + cr, s_scan, &tot_chunks);
    if (split) {
      BNode& nd = nodes[L.lb + j];
      const int k = carry_split + ks;
      const int li = L.le + 2 * k, ri = li + 1;
      if (ri < cap_nodes) {
        nd.left = li;
        nd.right = ri;
        BNode c;
        memset(&c, 0, sizeof(c));
        c.left = c.right = -1;
        c.level = nd.level + 1;
        c.has_prop = nd.has_prop;
        for (int q = 0; q < 3; ++q) c.prop[q] = nd.prop[q];
        c.start = nd.start;
        c.end = nd.start + nl;
        nodes[li] = c;
        c.start = nd.start + nl;
        c.end = nd.end;
        nodes[ri] = c;
        cbeg_next[2 * k] = carry_chunks + kc;
        cbeg_next[2 * k + 1] = carry_chunks + kc + cl;
      } else {  // cannot happen (a tree of n points has at most 2 n - 1 nodes)
        nd.split = 0;
        st->overflow = 1;
      }
    }
    carry_split += tot_split;
    carry_chunks += tot_chunks;
  }
  __syncthreads();
  if (tid == 0) {
    const int S = carry_split;
    const bool ok = L.le + 2 * S <= cap_nodes && !st->overflow;
    cbeg_next[2 * S] = carry_chunks;
    *ctl_next = LevelCtl{L.le, L.le + 2 * S, carry_chunks, (S > 0 && ok) ? 1 : 0};
    if (level + 2 < kMaxLevels + 2) level_begin[level + 2] = L.le + 2 * S;
    if (S == 0 || !ok) st->last_level = level + 1;
  }
}

__global__ void k_scatter(LevelArgs a, const int* coff, float* qx, float* qy, float* qz, double* qt) {
  const LevelCtl L = *a.ctl;
  const int c = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = threadIdx.x & 63;
  if (!L.active || c >= L.n_chunks) return;
  const ChunkRef r = chunk_ref(a, L, c);
  const BNode& nd = a.nodes[L.lb + r.j];
  const bool valid = lane < r.cnt;
  const int i = r.p0 + min(lane, r.cnt - 1);
  const float x = a.px[i], y = a.py[i], z = a.pz[i];
  const double t = a.pt[i];
  int o = i;
  if (nd.split) {
    // stable partition: a left point's rank is the number of left points
    // before it in the chunk, a right point's the number of right points
    const bool left = valid && tg::goes_left(nd.dir, nd.mean, x, y, z);
    const unsigned long long m = __ballot(left);
    const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
    const int rank_l = __popcll(m & below);
    const int cb = a.cbeg[r.j];
    const int before_left = coff[c] - coff[cb];
    const int before_right = (c - cb) * kC - before_left;
    o = left ? nd.start + before_left + rank_l : nd.start + nd.n_left + before_right + (lane - rank_l);
  }
  if (valid) {
    qx[o] = x;
    qy[o] = y;
    qz[o] = z;
    qt[o] = t;
  }
}

// ---- prune and breadth-first numbering -------------------------------------------

__global__ __launch_bounds__(kLevelThreads) void k_prune(BNode* nodes, const int* level_begin, int n_levels,
                                                         Status* st) {
  __shared__ int s_scan[16];
  const int tid = threadIdx.x;
  if (st->overflow) return;
  // keep flags, deepest level first (TreeCPU_::_prune)
  for (int l = n_levels - 1; l >= 0; --l) {
    const int b = level_begin[l], e = level_begin[l + 1];
    for (int i = b + tid; i < e; i += kLevelThreads) {
      BNode& nd = nodes[i];
      if (!nd.split)
        nd.keep = tg::good_leaf(nd.mean, nd.dir) ? 1 : 0;
      else
        nd.keep = (nodes[nd.left].keep | nodes[nd.right].keep) ? 1 : 0;
    }
    __syncthreads();
  }
  // the level order of the unpruned tree, restricted to the kept nodes, is
  // the breadth-first order of the pruned tree
  const int n = level_begin[n_levels];
  int carry = 0, carry_leaf = 0;
  for (int b = 0; b < n; b += kLevelThreads) {
    const int i = b + tid;
    int keep = 0, leaf = 0;
    if (i < n) {
      keep = nodes[i].keep;
      leaf = keep && !nodes[i].split;
    }
    int tot, tot_leaf;
    const int ex = block_exscan(keep, s_scan, &tot);
    const int exl = block_exscan(leaf, s_scan, &tot_leaf);
    if (i < n) {
      nodes[i].new_index = keep ? carry + ex : -1;
      nodes[i].leaf_index = leaf ? carry_leaf + exl : -1;
    }
    carry += tot;
    carry_leaf += tot_leaf;
  }
  if (tid == 0) {
    st->n_built = n;
    st->n_out = carry;
    st->n_leaves_out = carry_leaf;
    st->root_kept = n > 0 ? nodes[0].keep : 0;
  }
}

struct EmitArgs {
  const BNode* nodes;
  int n_nodes;
  float* dnodes;       // 8 floats per node
  float* dleaves;      // 8 floats per leaf
  float* ddts;         // per node
  kdh_node3f* full;    // Node_ records
  uint64_t* leaf_ids;  // leaf list
};

__global__ void k_emit(EmitArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n_nodes) return;
  const BNode& nd = a.nodes[i];
  if (!nd.keep) return;
  const int o = nd.new_index;
  int l = -1, r = -1;
  if (nd.split) {
    if (a.nodes[nd.left].keep) l = a.nodes[nd.left].new_index;
    if (a.nodes[nd.right].keep) r = a.nodes[nd.right].new_index;
  }
  float* rec = a.dnodes + (size_t)o * 8;
  rec[0] = nd.mean[0];
  rec[1] = nd.mean[1];
  rec[2] = nd.mean[2];
  rec[3] = nd.dir[0];
  rec[4] = nd.dir[1];
  rec[5] = nd.dir[2];
  rec[6] = __int_as_float(l);
  rec[7] = __int_as_float(r);
  a.ddts[o] = (float)nd.dts_mean;
  kdh_node3f f;
  f.dts_mean = nd.dts_mean;
  f.dts_covariance = nd.dts_cov;
  for (int k = 0; k < 3; ++k) {
    f.mean[k] = nd.mean[k];
    f.direction[k] = nd.dir[k];
  }
  f.idx_start = (uint64_t)nd.start;
  f.idx_end = (uint64_t)nd.end;
  f.idx_left = l;
  f.idx_right = r;
  a.full[o] = f;
  if (nd.leaf_index >= 0) {
    float* lr = a.dleaves + (size_t)nd.leaf_index * 8;
    lr[0] = nd.mean[0];
    lr[1] = nd.mean[1];
    lr[2] = nd.mean[2];
    lr[3] = nd.dir[0];
    lr[4] = nd.dir[1];
    lr[5] = nd.dir[2];
    lr[6] = (float)nd.dts_mean;
    lr[7] = __int_as_float(o);
    a.leaf_ids[nd.leaf_index] = (uint64_t)o;
  }
}

__global__ void k_points_out(const float* px, const float* py, const float* pz, const double* pt, int n, float* xyz,
                             double* stamps) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  xyz[3 * i] = px[i];
  xyz[3 * i + 1] = py[i];
  xyz[3 * i + 2] = pz[i];
  stamps[i] = pt[i];
}

size_t a256(size_t x) { return (x + 255) & ~(size_t)255; }

int ensure_ws(kdh_ctx* ctx, int64_t n) {
  Ws*& ws = reinterpret_cast<Ws*&>(ctx->tree_build_ws);
  if (ws && ws->cap >= n) return KDH_OK;
  int64_t cap = std::max<int64_t>(n, 4096);
  if (ws) cap = std::max<int64_t>(cap, 2 * ws->cap);
  kdh_tree_build_ws_free(ctx);
  const size_t N = (size_t)cap;
  const size_t bytes = 3 * 2 * a256(N * 4) + 2 * a256(N * 8) + a256(2 * N * sizeof(BNode)) + a256(N * 6 * 8) +
                       2 * a256(N * 4) + 2 * a256((N + 1) * 4) + a256((kMaxLevels + 2) * 4) +
                       a256(2 * sizeof(LevelCtl)) + a256(sizeof(Status)) + a256(3 * N * 4) + a256(N * 8);
  void* mem = nullptr;
  hipError_t e = hipMalloc(&mem, bytes);
  if (e != hipSuccess) return kdh_fail(ctx, KDH_ERR_NOMEM, std::string("kdh_tree_build: workspace: ") + hipGetErrorString(e));
  ws = new Ws();
  ws->cap = cap;
  ws->mem = mem;
  char* p = (char*)mem;
  auto take = [&](size_t b) {
    char* q = p;
    p += a256(b);
    return q;
  };
  for (int k = 0; k < 2; ++k) {
    ws->px[k] = (float*)take(N * 4);
    ws->py[k] = (float*)take(N * 4);
    ws->pz[k] = (float*)take(N * 4);
    ws->pt[k] = (double*)take(N * 8);
  }
  ws->nodes = (BNode*)take(2 * N * sizeof(BNode));
  ws->csum = (double*)take(N * 6 * 8);
  ws->cleft = (int*)take(N * 4);
  ws->coff = (int*)take(N * 4);
  ws->cbeg[0] = (int*)take((N + 1) * 4);
  ws->cbeg[1] = (int*)take((N + 1) * 4);
  ws->level_begin = (int*)take((kMaxLevels + 2) * 4);
  ws->ctl = (LevelCtl*)take(2 * sizeof(LevelCtl));
  ws->status = (Status*)take(sizeof(Status));
  ws->stage_xyz = (float*)take(3 * N * 4);
  ws->stage_t = (double*)take(N * 8);
  e = hipHostMalloc((void**)&ws->h_status, sizeof(Status) * (kMaxLevels / kBatchLevels + 2), hipHostMallocDefault);
  if (e != hipSuccess) {
    kdh_tree_build_ws_free(ctx);
    return kdh_fail(ctx, KDH_ERR_NOMEM, "kdh_tree_build: pinned status");
  }
  return KDH_OK;
}

}  // namespace

void kdh_tree_build_ws_free(kdh_ctx* ctx) {
  Ws*& ws = reinterpret_cast<Ws*&>(ctx->tree_build_ws);
  if (!ws) return;
  if (ws->mem) hipFree(ws->mem);
  if (ws->h_status) hipHostFree(ws->h_status);
  delete ws;
  ws = nullptr;
}

extern "C" {

int kdh_tree_build(kdh_ctx* ctx, const float* xyz, const double* stamps, int64_t n_points,
                   const kdh_tree_gen_params* params, int inputs_on_device, kdh_tree** out) {
  if (!ctx || !out || !params) return KDH_ERR_INVALID;
  *out = nullptr;
  if (n_points < 0 || (n_points > 0 && !xyz)) return kdh_fail(ctx, KDH_ERR_INVALID, "kdh_tree_build: bad points");
  if (n_points > (int64_t)(1 << 28))
    return kdh_fail(ctx, KDH_ERR_INVALID, "kdh_tree_build: more than 2^28 points");
  const kdh_tree_gen_params P = *params;
  if (n_points < (int64_t)P.leaf_min_points || n_points == 0) return KDH_OK;  // makeTree returns nullptr
  std::lock_guard<std::mutex> lk(ctx->mu);
  KDH_HIP_TRY(ctx, hipSetDevice(ctx->device));
  int rc = ensure_ws(ctx, n_points);
  if (rc) return rc;
  Ws* ws = reinterpret_cast<Ws*>(ctx->tree_build_ws);
  const int n = (int)n_points;
  const int cap_nodes = 2 * (int)ws->cap;
  hipStream_t s = ctx->stream;

  // points into the first buffer, structure of arrays
  const float* d_xyz = xyz;
  const double* d_t = stamps;
  if (!inputs_on_device) {
    KDH_HIP_TRY(ctx, hipMemcpyAsync(ws->stage_xyz, xyz, (size_t)n * 12, hipMemcpyHostToDevice, s));
    d_xyz = ws->stage_xyz;
    if (stamps) {
      KDH_HIP_TRY(ctx, hipMemcpyAsync(ws->stage_t, stamps, (size_t)n * 8, hipMemcpyHostToDevice, s));
      d_t = ws->stage_t;
    }
  }
  hipLaunchKernelGGL(k_stage, dim3((n + 255) / 256), dim3(256), 0, s, d_xyz, d_t, n, ws->px[0], ws->py[0], ws->pz[0],
                     ws->pt[0]);
  hipLaunchKernelGGL(k_init, dim3(1), dim3(kLevelThreads), 0, s, ws->pt[0], n, ws->nodes, ws->ctl, ws->cbeg[0],
                     ws->level_begin, ws->status);
  KDH_HIP_TRY(ctx, hipGetLastError());

  const int n_chunks_max = (n + kC - 1) / kC;
  auto enqueue_level = [&](int L) -> hipError_t {
    const int cur = L & 1, nxt = cur ^ 1;
    const int64_t width_bound = std::min<int64_t>((int64_t)1 << std::min(L, 30), n);
    const int64_t chunk_bound = std::min<int64_t>(n, n_chunks_max + width_bound);
    const int gc = (int)((chunk_bound + 3) / 4);  // one wave per chunk
    const int gw = (int)std::min<int64_t>((width_bound + 3) / 4, 4096);
    LevelArgs a;
    a.nodes = ws->nodes;
    a.ctl = ws->ctl + cur;
    a.cbeg = ws->cbeg[cur];
    a.px = ws->px[cur];
    a.py = ws->py[cur];
    a.pz = ws->pz[cur];
    a.pt = ws->pt[cur];
    a.csum = ws->csum;
    a.st = ws->status;
    hipLaunchKernelGGL(k_mom1, dim3(gc), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_red1, dim3(gw), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_mom2, dim3(gc), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_red2, dim3(gw), dim3(256), 0, s, a, P);
    hipLaunchKernelGGL(k_count, dim3(gc), dim3(256), 0, s, a, ws->cleft);
    hipLaunchKernelGGL(k_level, dim3(1), dim3(kLevelThreads), 0, s, ws->nodes, ws->ctl + cur, ws->ctl + nxt,
                       ws->cbeg[cur], ws->cbeg[nxt], ws->cleft, ws->coff, ws->level_begin, L, cap_nodes, ws->status);
    hipLaunchKernelGGL(k_scatter, dim3(gc), dim3(256), 0, s, a, ws->coff, ws->px[nxt], ws->py[nxt], ws->pz[nxt],
                       ws->pt[nxt]);
    return hipGetLastError();
  };

  // Levels are queued one batch ahead of the status the host has seen; the
  // levels queued past the last one find no nodes and return at once.
  int queued = 0, n_levels = -1;
  std::vector<hipEvent_t> evs;
  auto release = [&]() {
    for (hipEvent_t ev : evs) hipEventDestroy(ev);
  };
  while (n_levels < 0) {
    if (queued >= kMaxLevels) {
      release();
      return kdh_fail(ctx, KDH_ERR_INVALID, "kdh_tree_build: deeper than 4096 levels");
    }
    hipError_t e = hipSuccess;
    for (int k = 0; k < kBatchLevels && e == hipSuccess; ++k) e = enqueue_level(queued++);
    hipEvent_t ev = nullptr;
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    if (e == hipSuccess) evs.push_back(ev);
    const int b = (int)evs.size() - 1;
    if (e == hipSuccess)
      e = hipMemcpyAsync(ws->h_status + b, ws->status, sizeof(Status), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipEventRecord(ev, s);
    if (e == hipSuccess && b >= 1) {
      e = hipEventSynchronize(evs[b - 1]);
      if (e == hipSuccess && ws->h_status[b - 1].last_level >= 0) n_levels = ws->h_status[b - 1].last_level;
    }
    if (e != hipSuccess) {
      release();
      return kdh_hip_fail(ctx, e, "kdh_tree_build: level queue");
    }
  }
  release();

  hipLaunchKernelGGL(k_prune, dim3(1), dim3(kLevelThreads), 0, s, ws->nodes, ws->level_begin, n_levels, ws->status);
  KDH_HIP_TRY(ctx, hipGetLastError());
  Status st;
  KDH_HIP_TRY(ctx, hipMemcpyAsync(ws->h_status, ws->status, sizeof(Status), hipMemcpyDeviceToHost, s));
  KDH_HIP_TRY(ctx, hipStreamSynchronize(s));
  st = ws->h_status[0];
  if (st.overflow) return kdh_fail(ctx, KDH_ERR_INVALID, "kdh_tree_build: node capacity exceeded");
  const int n_out = st.root_kept ? st.n_out : 0, n_leaves = st.root_kept ? st.n_leaves_out : 0;
  if (n_leaves < P.tree_min_good_leaves) return KDH_OK;  // makeTree returns nullptr

  // the tree: the kdh_tree_upload layout, then the full Node_ records, the
  // leaf list and the permuted points
  const size_t node_bytes = a256((size_t)std::max(n_out, 1) * 32);
  const size_t leaf_bytes = a256((size_t)std::max(n_leaves, 1) * 32);
  const size_t dts_bytes = a256((size_t)std::max(n_out, 1) * 4);
  const size_t full_bytes = a256((size_t)std::max(n_out, 1) * sizeof(kdh_node3f));
  const size_t ids_bytes = a256((size_t)std::max(n_leaves, 1) * 8);
  const size_t xyz_bytes = a256((size_t)n * 12);
  const size_t t_bytes = a256((size_t)n * 8);
  const size_t total = node_bytes + leaf_bytes + dts_bytes + full_bytes + ids_bytes + xyz_bytes + t_bytes;
  void* d = nullptr;
  hipError_t e = hipMalloc(&d, total);
  if (e != hipSuccess) return kdh_fail(ctx, KDH_ERR_NOMEM, std::string("kdh_tree_build: hipMalloc: ") + hipGetErrorString(e));
  char* p = (char*)d;
  EmitArgs ea;
  ea.nodes = ws->nodes;
  ea.n_nodes = n_out > 0 ? st.n_built : 0;
  ea.dnodes = (float*)p;
  ea.dleaves = (float*)(p + node_bytes);
  ea.ddts = (float*)(p + node_bytes + leaf_bytes);
  ea.full = (kdh_node3f*)(p + node_bytes + leaf_bytes + dts_bytes);
  ea.leaf_ids = (uint64_t*)(p + node_bytes + leaf_bytes + dts_bytes + full_bytes);
  float* out_xyz = (float*)(p + node_bytes + leaf_bytes + dts_bytes + full_bytes + ids_bytes);
  double* out_t = (double*)(p + node_bytes + leaf_bytes + dts_bytes + full_bytes + ids_bytes + xyz_bytes);
  if (ea.n_nodes > 0) hipLaunchKernelGGL(k_emit, dim3((ea.n_nodes + 255) / 256), dim3(256), 0, s, ea);
  hipLaunchKernelGGL(k_points_out, dim3((n + 255) / 256), dim3(256), 0, s, ws->px[0], ws->py[0], ws->pz[0],
                     ws->pt[0], n, out_xyz, out_t);
  e = hipGetLastError();
  kdh_node3f root{};
  if (e == hipSuccess && n_out > 0) e = hipMemcpyAsync(&root, ea.full, sizeof(root), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    hipFree(d);
    return kdh_hip_fail(ctx, e, "kdh_tree_build: emit");
  }
  kdh_tree* t = new kdh_tree();
  t->device = ctx->device;
  t->dmem = d;
  t->bytes = total;
  t->dt.nodes = (const float4*)ea.dnodes;
  t->dt.leaves = (const float4*)ea.dleaves;
  t->node_dts = ea.ddts;
  t->dt.node_dts = ea.ddts;
  t->dt.n_nodes = n_out;
  t->dt.n_leaves = n_leaves;
  t->root_mean = n_out > 0 ? v3(root.mean[0], root.mean[1], root.mean[2]) : v3(0.0f, 0.0f, 0.0f);
  t->full_nodes = ea.full;
  t->full_leaves = ea.leaf_ids;
  t->full_xyz = out_xyz;
  t->full_stamps = out_t;
  t->n_points = n;
  *out = t;
  return KDH_OK;
}

int kdh_tree_download(kdh_ctx* ctx, const kdh_tree* tree, kdh_node3f* nodes, uint64_t* leaves, float* xyz,
                      double* stamps) {
  if (!ctx || !tree) return KDH_ERR_INVALID;
  if (!tree->full_nodes)
    return kdh_fail(ctx, KDH_ERR_INVALID, "kdh_tree_download: only trees built by kdh_tree_build keep full records");
  std::lock_guard<std::mutex> lk(ctx->mu);
  KDH_HIP_TRY(ctx, hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  if (nodes && tree->dt.n_nodes > 0)
    KDH_HIP_TRY(ctx, hipMemcpyAsync(nodes, tree->full_nodes, sizeof(kdh_node3f) * tree->dt.n_nodes,
                                    hipMemcpyDeviceToHost, s));
  if (leaves && tree->dt.n_leaves > 0)
    KDH_HIP_TRY(ctx, hipMemcpyAsync(leaves, tree->full_leaves, 8 * (size_t)tree->dt.n_leaves, hipMemcpyDeviceToHost, s));
  if (xyz && tree->n_points > 0)
    KDH_HIP_TRY(ctx, hipMemcpyAsync(xyz, tree->full_xyz, 12 * (size_t)tree->n_points, hipMemcpyDeviceToHost, s));
  if (stamps && tree->n_points > 0)
    KDH_HIP_TRY(ctx, hipMemcpyAsync(stamps, tree->full_stamps, 8 * (size_t)tree->n_points, hipMemcpyDeviceToHost, s));
  KDH_HIP_TRY(ctx, hipStreamSynchronize(s));
  return KDH_OK;
}

}  // extern "C"
