// PCA KD-tree generation on the device (kdh_tree_build): TreeGenerator_::
// makeTree (tree/tree_generator_impl_.h:9-19) followed by the device upload
// (Tree_CUDAfromCPU_, tree/tree_cuda_impl_.h:198-222), without a host round
// trip.  The tree is bit-identical to the host builder's (tree_build.cpp):
// both evaluate the shared labelling, eigen-solver and split rule of
// kdh_treegen.h, sum node moments in its canonical order and keep the point
// order of each side of a split.
//
// The build runs level by level.  Level L holds a contiguous range of nodes;
// their points are cut into chunks of 64 (a chunk never straddles two nodes;
// an explicit chunk-to-node map replaces any search).  Per level:
//   mom1, mom2  wave per chunk: pairwise sums of (x, y, z, dts, dts^2), then
//               of the products centred on the FP64 mean
//   red1, red2  wave per node: the chunk sums in the canonical pairwise order;
//               red2 also solves the eigenproblem and labels the node
//   count       wave per chunk: points going left (ballot); integer totals
//               per node through 64-bit atomics whose return value tells
//               the last chunk, which decides the split
//   alloc       prefix sums over 1024-node blocks of the split flags and the
//               children's chunks; alloc2 (one workgroup) the block offsets
//   split       wave per chunk: stable partition into the other point buffer
//               and the next level's chunk-to-node map; thread per node: the
//               children
// Floating-point partial sums cross workgroups only at kernel boundaries
// (the XCDs' L2 caches are not coherent with each other; a fence that
// publishes data across them writes the L2 back).  Integer totals and the
// prune's keep flags travel in atomic return values, which need no fence.
// The node and chunk counts stay on the device; the host queues levels ahead
// and reads a status word every few levels.  The prune climbs from the leaves
// (the second child to arrive at a parent decides it), the kept nodes are
// numbered by prefix sums over the level order, which is the breadth-first
// order of the pruned tree, and a last pass writes the device tree layout of
// kdh_tree_upload plus the full Node_ records, the leaf list and the permuted
// points.
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
constexpr int kScanThreads = 1024;  // prefix-sum blocks
constexpr int kMaxLevels = 4096;
constexpr int kBatchLevels = 4;     // levels queued between status checks
constexpr int kSubLimit = 2048;     // subtrees of at most this many points finish in one workgroup
constexpr int kSubThreads = 256;

// One node of the level-ordered build (before pruning).
struct BNode {
  int start, end;     // point range
  int left, right;    // children (-1)
  int parent;         // -1 for the root
  int level;
  int is_node;        // the labelling asks for a split
  int split;          // split into two non-empty sides
  int has_prop;       // FlatLeafPolicy_ propagated normal (inherited, then own)
  int n_left;
  int ex_split;       // within its prefix-sum block: splits before it
  int ex_chunks;      //   children's chunks before it
  int keep;           // survives the prune
  int ex_keep;        // within its prefix-sum block: kept nodes before it
  int ex_leaf;        //   kept leaves before it
  int gidx;           // subtree phase: the global index of a staged node
  int n_depths;       // subtree phase, staged root: depths of its subtree
  int n_staged;       //   and nodes
  float mean[3], dir[3], evec0[3], prop[3];
  double dts_mean, dts_cov;
  double s1[5];       // (x, y, z, dts, dts^2) sums
  double c2[6];       // subtree phase: centred sums waiting for the labelling
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
  int handoff;        // 1: the levels from handoff_level on are built by subtrees
  int handoff_level;
  int handoff_lb, handoff_n;  // the subtree roots: nodes [lb, lb + n)
  int max_depth;      // deepest subtree (levels below its root, root included)
  int n_built;
  int pad;
  double ts_start;
};

struct Ws {
  int64_t cap = 0;    // points
  float* px[2] = {nullptr, nullptr};
  float* py[2] = {nullptr, nullptr};
  float* pz[2] = {nullptr, nullptr};
  double* pt[2] = {nullptr, nullptr};
  BNode* nodes = nullptr;   // 2 cap
  double* csum = nullptr;   // chunk sums, 6 per chunk (cap)
  double* gsum = nullptr;   // group sums of the top levels, 6 per slot (cap)
  int* cleft = nullptr;     // left points per chunk (cap)
  int* gleft = nullptr;     // left points per 64-chunk group slot (cap)
  unsigned long long* cnt64 = nullptr;  // per node: arrivals << 32 | count (2 cap); zero between uses
  unsigned long long* cnt64_grp = nullptr;  // per 64-chunk group slot, the same (cap); zero between uses
  int* keep_word = nullptr; // per node: 2 x arrivals + keep bits (2 cap); zero between uses
  int* chunk_node[2] = {nullptr, nullptr};  // node of each chunk (cap each)
  int* blk_a = nullptr;     // prefix-sum block totals, then offsets (n_blk)
  int* blk_b = nullptr;
  int* blk_c = nullptr;     // largest child per block
  int n_blk = 0;
  BNode* stage = nullptr;   // subtree phase: subtree s stages its nodes at [2 start_s, 2 end_s) (2 cap)
  int* depth_cnt = nullptr; // subtree phase: subtree s's nodes per depth at [start_s + s, ...) (2 cap)
  int* cbeg[2] = {nullptr, nullptr};  // cap + 1 each
  int* level_begin = nullptr;         // kMaxLevels + 2
  LevelCtl* ctl = nullptr;            // [2]
  Status* status = nullptr;
  unsigned long long* ts_key = nullptr;  // smallest stamp, order key; ~0 between builds
  float* stage_xyz = nullptr;         // 3 cap (host inputs)
  double* stage_t = nullptr;          // cap
  Status* h_status = nullptr;         // pinned
  int* h_count = nullptr;             // pinned
  void* mem = nullptr;
  // the first batch of levels as a graph; its launches depend only on the
  // workspace, the parameters and the batch length
  hipGraphExec_t first_exec = nullptr;
  int first_levels = 0;
  // two more streams (default priority) for the subtree size classes, with
  // the fork and join events; made at first use
  hipStream_t aux[2] = {nullptr, nullptr};
  hipEvent_t fork = nullptr, join[2] = {nullptr, nullptr};
  kdh_tree_gen_params first_params{};
};

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

// Exclusive prefix sum over the threads of the block (a multiple of 64, at
// most 1024); *total gets the block sum.  s: 16 ints of LDS.
__device__ int block_exscan(int v, int* s, int* total) {
  const int lane = lane_id(), wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) s[wid] = x;
  __syncthreads();
  if (wid == 0) {
    int t = lane < nw ? s[lane] : 0;
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) {
      const int y = __shfl_up(t, o, 64);
      if (lane >= o) t += y;
    }
    if (lane < nw) s[lane] = t;
  }
  __syncthreads();
  const int off = wid ? s[wid - 1] : 0;
  *total = s[nw - 1];
  __syncthreads();
  return off + x - v;
}

__device__ __forceinline__ int wave_sum_int(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Lane i gets x of lane i + S within its row of 16 lanes (DPP row_shl:S on
// both halves; a VALU move, no LDS round trip).  Lanes whose source lies past
// the row end get 0.
template <int S>
__device__ __forceinline__ double row_down(double x) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), 0x100 + S, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), 0x100 + S, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// x of lane src (wave-uniform): a scalar read of that lane's register.
__device__ __forceinline__ double lane_value(double x, int src) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(x), src);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(x), src);
  return __hiloint2double(hi, lo);
}

// One pairwise round of the canonical order: lane i (a multiple of 2S) adds
// lane i + S when i + S < cnt.
template <int S, int K>
__device__ __forceinline__ void pairwise_round(double (&v)[K], int cnt, int lane) {
  if (S >= cnt) return;  // no lane takes (cnt is wave-uniform)
  const bool take = ((lane & (2 * S - 1)) == 0) && (lane + S < cnt);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const double o = row_down<S>(v[k]);
    if (take) v[k] += o;
  }
}

// Rounds of value stride 4S (S = 1, 2, 4, 8) over chunks held four values
// per lane in rows of 16 (chunk_sums_rows): lane j of a row (a multiple of
// 2S) adds lane j + S when value 4(j + S) exists.
template <int S, int K>
__device__ __forceinline__ void row_round(double (&v)[K], int cnt, int j) {
  const bool take = ((j & (2 * S - 1)) == 0) && (4 * (j + S) < cnt);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const double o = row_down<S>(v[k]);
    if (take) v[k] += o;
  }
}

// One group of the canonical order in a wave: lane i holds value i (i < cnt);
// every lane ends with the group sum.  The rounds of stride 1..8 stay inside
// rows of 16 lanes (DPP moves); the last two, (r0 + r16) + (r32 + r48) over
// the row sums, are formed from scalar reads -- the same additions in the
// same order as the stride-doubling tree.
template <int K>
__device__ __forceinline__ void wave_pairwise(double (&v)[K], int cnt) {
  const int lane = lane_id();
  pairwise_round<1>(v, cnt, lane);
  pairwise_round<2>(v, cnt, lane);
  pairwise_round<4>(v, cnt, lane);
  pairwise_round<8>(v, cnt, lane);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const double r0 = lane_value(v[k], 0);
    if (cnt <= 16) {
      v[k] = r0;
      continue;
    }
    const double r16 = lane_value(v[k], 16), r32 = lane_value(v[k], 32), r48 = lane_value(v[k], 48);
    const double a = r0 + r16;                 // 16 < cnt
    const double b = 48 < cnt ? r32 + r48 : r32;
    v[k] = 32 < cnt ? a + b : a;
  }
}

// The canonical reduction of the m values at vals[base..] (6 doubles apart)
// by one wave, in place; every lane ends with the result.  kRedBatch groups
// are loaded per round trip; group g's sum goes to slot g, which lies in a
// group already consumed.
constexpr int kRedBatch = 8;
template <int K>
__device__ void wave_reduce_values(double* vals, int base, int m, double (&out)[K]) {
  const int lane = lane_id();
  while (m > 1) {
    const int groups = (m + tg::kGroup - 1) / tg::kGroup;
    for (int g = 0; g < groups; g += kRedBatch) {
      double v[kRedBatch][K];
#pragma unroll
      for (int u = 0; u < kRedBatch; ++u) {
        const int b = (g + u) * tg::kGroup;
        const int cnt = g + u < groups ? min(m - b, tg::kGroup) : 1;
        const double* src = vals + (size_t)(base + (g + u < groups ? b + min(lane, cnt - 1) : 0)) * 6;
#pragma unroll
        for (int k = 0; k < K; ++k) v[u][k] = src[k];
      }
#pragma unroll
      for (int u = 0; u < kRedBatch; ++u) {
        if (g + u >= groups) break;
        const int b = (g + u) * tg::kGroup;
        wave_pairwise<K>(v[u], min(m - b, tg::kGroup));
      }
      if (lane == 0) {
#pragma unroll
        for (int u = 0; u < kRedBatch; ++u) {
          if (g + u >= groups) break;
          double* dst = vals + (size_t)(base + g + u) * 6;
#pragma unroll
          for (int k = 0; k < K; ++k) dst[k] = v[u][k];
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
    m = groups;
  }
  const double* r = vals + (size_t)base * 6;
#pragma unroll
  for (int k = 0; k < K; ++k) out[k] = r[k];
}

// ---- set-up ------------------------------------------------------------------

// Order-preserving integer image of a double (not NaN) and back.
__device__ __forceinline__ unsigned long long order_key(double d) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(d);
  return (b >> 63) ? ~b : (b | (1ull << 63));
}
__device__ __forceinline__ double from_order_key(unsigned long long k) {
  return __longlong_as_double((long long)((k >> 63) ? (k & ~(1ull << 63)) : ~k));
}

// Points into structure-of-arrays form.  The smallest stamp (fmin over the
// cloud from DBL_MAX, absent stamps counting as 0) goes into *ts_key as an
// order-preserving integer, one atomicMin per block: the minimum does not
// depend on the order of arrival.  256 threads per block.
__global__ __launch_bounds__(256) void k_stage(const float* __restrict__ xyz, const double* __restrict__ stamps,
                                               int n, float* px, float* py, float* pz, double* pt,
                                               unsigned long long* ts_key) {
  __shared__ double s_min[4];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double m = 1.7976931348623157e308;
  if (i < n) {
    px[i] = xyz[3 * i];
    py[i] = xyz[3 * i + 1];
    pz[i] = xyz[3 * i + 2];
    const double t = stamps ? stamps[i] : 0.0;
    pt[i] = t;
    m = fmin(m, t);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmin(m, __shfl_xor(m, o, 64));
  if (lane_id() == 0) s_min[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    // most blocks find the minimum already below theirs (a stale read only
    // costs an atomic that changes nothing)
    const unsigned long long key = order_key(fmin(fmin(s_min[0], s_min[1]), fmin(s_min[2], s_min[3])));
    if (key < __atomic_load_n(ts_key, __ATOMIC_RELAXED)) atomicMin(ts_key, key);
  }
}

// The root, the first level's control and chunk map, and the status; the
// stamp minimum is read from (and ts_key reset for the next build).
__global__ __launch_bounds__(kScanThreads) void k_init(unsigned long long* ts_key, int n, BNode* nodes,
                                                       LevelCtl* ctl, int* cbeg0, int* chunk_node0, int* level_begin,
                                                       Status* st) {
  const int nc0 = (n + kC - 1) / kC;
  for (int c = threadIdx.x; c < nc0; c += kScanThreads) chunk_node0[c] = 0;
  if (threadIdx.x == 0) {
    BNode r;
    memset(&r, 0, sizeof(r));
    r.start = 0;
    r.end = n;
    r.left = r.right = r.parent = -1;
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
    s.ts_start = from_order_key(*ts_key);
    *st = s;
    *ts_key = ~0ull;
  }
}

// ---- per level ---------------------------------------------------------------

struct LevelArgs {
  BNode* nodes;
  const LevelCtl* ctl;  // this level's
  LevelCtl* ctl_next;
  const int* cbeg;      // this level's chunk starts (w + 1)
  int* cbeg_next;
  const float *px, *py, *pz;
  const double* pt;
  float *qx, *qy, *qz;  // the other point buffer
  double* qt;
  const int* chunk_node; // this level's
  int* chunk_node_next;
  double* csum;
  double* gsum;         // first-round group sums of the top levels (k_gsum)
  int* cleft;
  int* gleft;
  unsigned long long* cnt64;
  unsigned long long* cnt64_grp;
  int* blk_a;
  int* blk_b;
  int* blk_c;
  int* level_begin;
  Status* st;
  int level;
  int cap_nodes;
};

// The chunk of this wave.  Wave-uniform.
struct ChunkRef {
  int j;       // node within the level
  int cb, m;   // the node's first chunk and chunk count
  int p0;      // first point
  int cnt;     // points in the chunk (1..64)
};

__device__ __forceinline__ ChunkRef chunk_ref(const LevelArgs& a, const LevelCtl& L, int c) {
  ChunkRef r;
  r.j = a.chunk_node[c] - L.lb;
  r.cb = a.cbeg[r.j];
  r.m = a.cbeg[r.j + 1] - r.cb;
  const BNode& nd = a.nodes[L.lb + r.j];
  r.p0 = nd.start + (c - r.cb) * kC;
  r.cnt = min(kC, nd.end - r.p0);
  return r;
}

__device__ __forceinline__ void store_chunk(double* csum, int c, const double* v, int K) {
  if (lane_id() == 0)
    for (int k = 0; k < K; ++k) csum[(size_t)c * 6 + k] = v[k];
}

__global__ void k_mom1(LevelArgs a) {
  const LevelCtl L = *a.ctl;
  const int c = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = lane_id();
  if (!L.active || c >= L.n_chunks) return;
  const ChunkRef r = chunk_ref(a, L, c);
  const int i = r.p0 + min(lane, r.cnt - 1);
  const double d = a.pt[i] - a.st->ts_start;
  double v[5] = {(double)a.px[i], (double)a.py[i], (double)a.pz[i], d, d * d};
  wave_pairwise<5>(v, r.cnt);
  store_chunk(a.csum, c, v, 5);
}

__global__ void k_mom2(LevelArgs a) {
  const LevelCtl L = *a.ctl;
  const int c = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = lane_id();
  if (!L.active || c >= L.n_chunks) return;
  const ChunkRef r = chunk_ref(a, L, c);
  const BNode& nd = a.nodes[L.lb + r.j];
  const double inv = 1.0 / (double)(nd.end - nd.start);
  const double mu0 = nd.s1[0] * inv, mu1 = nd.s1[1] * inv, mu2 = nd.s1[2] * inv;
  const int i = r.p0 + min(lane, r.cnt - 1);
  const double d0 = (double)a.px[i] - mu0, d1 = (double)a.py[i] - mu1, d2 = (double)a.pz[i] - mu2;
  double v[6] = {d0 * d0, d0 * d1, d0 * d2, d1 * d1, d1 * d2, d2 * d2};
  wave_pairwise<6>(v, r.cnt);
  store_chunk(a.csum, c, v, 6);
}

// The first round of the nodes' canonical reductions spread over many waves:
// group g of node j (its chunks 64g .. 64g + 63) is summed by one wave into
// gsum[cbeg[j] + g].  Used on the top levels, where few nodes hold many
// chunks each; k_red1 / k_red2 then reduce the group sums.  Grid: one
// workgroup column per node, gridDim.y columns of waves share its groups.
template <int K>
__global__ __launch_bounds__(1024) void k_gsum(LevelArgs a) {
  const LevelCtl L = *a.ctl;
  const int j = blockIdx.x;
  if (!L.active || j >= L.le - L.lb) return;
  const int base = a.cbeg[j], m = a.cbeg[j + 1] - base;
  const int groups = (m + tg::kGroup - 1) / tg::kGroup;
  const int per = blockDim.x >> 6, lane = lane_id();
  for (int g = blockIdx.y * per + (threadIdx.x >> 6); g < groups; g += gridDim.y * per) {
    const int b = g * tg::kGroup, cnt = min(m - b, tg::kGroup);
    const double* src = a.csum + (size_t)(base + b + min(lane, cnt - 1)) * 6;
    double v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = src[k];
    wave_pairwise<K>(v, cnt);
    if (lane == 0)
      for (int k = 0; k < K; ++k) a.gsum[(size_t)(base + g) * 6 + k] = v[k];
  }
}

// grouped: the first round was done by k_gsum.
__global__ void k_red1(LevelArgs a, int grouped) {
  const LevelCtl L = *a.ctl;
  if (!L.active) return;
  const int w = L.le - L.lb;
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int n_waves = (gridDim.x * blockDim.x) >> 6;
  for (int j = wave; j < w; j += n_waves) {
    const int base = a.cbeg[j], m = a.cbeg[j + 1] - base;
    double s1[5];
    if (grouped)
      wave_reduce_values<5>(a.gsum, base, (m + tg::kGroup - 1) / tg::kGroup, s1);
    else
      wave_reduce_values<5>(a.csum, base, m, s1);
    if (lane_id() == 0) {
      BNode& nd = a.nodes[L.lb + j];
#pragma unroll
      for (int k = 0; k < 5; ++k) nd.s1[k] = s1[k];
    }
  }
}

__global__ void k_red2(LevelArgs a, kdh_tree_gen_params P, int grouped) {
  const LevelCtl L = *a.ctl;
  if (!L.active) return;
  const int w = L.le - L.lb;
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int n_waves = (gridDim.x * blockDim.x) >> 6;
  for (int j = wave; j < w; j += n_waves) {
    const int base = a.cbeg[j], m = a.cbeg[j + 1] - base;
    double c2[6];
    if (grouped)
      wave_reduce_values<6>(a.gsum, base, (m + tg::kGroup - 1) / tg::kGroup, c2);
    else
      wave_reduce_values<6>(a.csum, base, m, c2);
    if (lane_id() == 0) {
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

// Left points per chunk, per 64-chunk group and per node (integer sums: any
// order).  The node total and the arrival count share one 64-bit atomic, so
// the chunk that arrives last reads the total from its return value and
// decides the split.  Every other chunk of the node has evaluated its points
// before its atomic, so the one-sided case may then replace the direction.
__global__ void k_count(LevelArgs a) {
  const LevelCtl L = *a.ctl;
  const int c = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = lane_id();
  if (!L.active || c >= L.n_chunks) return;
  const ChunkRef r = chunk_ref(a, L, c);
  BNode& nd = a.nodes[L.lb + r.j];
  if (!nd.is_node) {
    if (c == r.cb && lane == 0) {
      nd.split = 0;
      nd.n_left = 0;
    }
    return;
  }
  bool left = false;
  if (lane < r.cnt) {
    const int i = r.p0 + lane;
    left = tg::goes_left(nd.dir, nd.mean, a.px[i], a.py[i], a.pz[i]);
  }
  const int cnt = __popcll(__ballot(left));
  if (lane != 0) return;
  int total = cnt;
  if (r.m > 1) {
    a.cleft[c] = cnt;
    // at most 64 chunks meet on one word: first per 64-chunk group, then the
    // last chunk of each group carries the group's total to the node's word
    const int g = (c - r.cb) >> 6;
    const int gcnt = min(kC, r.m - kC * g);
    const int m2 = (r.m + kC - 1) / kC;
    unsigned long long* word = m2 > 1 ? &a.cnt64_grp[r.cb + g] : &a.cnt64[L.lb + r.j];
    unsigned long long old = atomicAdd(word, (1ull << 32) | (unsigned long long)cnt);
    if ((int)(old >> 32) != (m2 > 1 ? gcnt : r.m) - 1) return;
    total = (int)(old & 0xffffffffull) + cnt;
    *word = 0ull;
    if (m2 > 1) {
      a.gleft[r.cb + g] = total;
      old = atomicAdd(&a.cnt64[L.lb + r.j], (1ull << 32) | (unsigned long long)total);
      if ((int)(old >> 32) != m2 - 1) return;
      total = (int)(old & 0xffffffffull) + total;
      a.cnt64[L.lb + r.j] = 0ull;
    }
  }
  const int n = nd.end - nd.start;
  const int split = (total > 0 && total < n) ? 1 : 0;
  if (!split)  // one-sided split: a leaf with its normal
    for (int k = 0; k < 3; ++k) nd.dir[k] = nd.evec0[k];
  nd.split = split;
  nd.n_left = total;
}

// Children's slots in breadth-first order: prefix sums of the split flags
// and of the children's chunks within each 1024-node block.
// The next level's range and chunk count from the level's totals (S split
// nodes, `chunks` child chunks, the largest child), and the stop and
// hand-over decisions.  One thread.
__device__ void alloc_tail(const LevelArgs& a, const LevelCtl& L, int S, int chunks, int largest) {
  const bool ok = L.le + 2 * S <= a.cap_nodes;
  if (!ok) a.st->overflow = 1;
  // every child small enough: the subtree phase builds the rest
  const bool handoff = S > 0 && ok && largest <= kSubLimit;
  a.cbeg_next[2 * S] = chunks;
  *a.ctl_next = LevelCtl{L.le, L.le + 2 * S, chunks, (S > 0 && ok && !handoff) ? 1 : 0};
  if (a.level + 2 < kMaxLevels + 2) a.level_begin[a.level + 2] = L.le + 2 * S;
  if (handoff) {
    a.st->handoff = 1;
    a.st->handoff_level = a.level + 1;
    a.st->handoff_lb = L.le;
    a.st->handoff_n = 2 * S;
  }
  if (S == 0 || !ok || handoff) a.st->last_level = a.level + 1;
}

// Block-local prefix sums of the split flags and child chunk counts.  SINGLE
// (a level of at most kScanThreads nodes, one block): the block offsets are
// zero and this block also finishes the level (alloc_tail), which otherwise
// k_alloc2 does over the block totals.
template <bool SINGLE>
__global__ __launch_bounds__(kScanThreads) void k_alloc(LevelArgs a) {
  __shared__ int s_scan[16];
  const LevelCtl L = *a.ctl;
  if (!L.active) {
    if (SINGLE && threadIdx.x == 0) *a.ctl_next = LevelCtl{L.le, L.le, 0, 0};
    return;
  }
  const int w = L.le - L.lb;
  const int j = blockIdx.x * kScanThreads + threadIdx.x;
  __shared__ int s_max;
  if (threadIdx.x == 0) s_max = 0;
  int split = 0, chunks = 0, largest = 0;
  if (j < w) {
    const BNode& nd = a.nodes[L.lb + j];
    if (nd.split) {
      split = 1;
      const int n = nd.end - nd.start;
      chunks = (nd.n_left + kC - 1) / kC + (n - nd.n_left + kC - 1) / kC;
      largest = max(nd.n_left, n - nd.n_left);
    }
  }
  int tot_s, tot_c;
  const int es = block_exscan(split, s_scan, &tot_s);
  const int ec = block_exscan(chunks, s_scan, &tot_c);
  if (largest) atomicMax(&s_max, largest);
  __syncthreads();
  if (j < w) {
    BNode& nd = a.nodes[L.lb + j];
    nd.ex_split = es;
    nd.ex_chunks = ec;
  }
  if (threadIdx.x == 0) {
    if (SINGLE) {
      a.blk_a[0] = 0;
      a.blk_b[0] = 0;
      alloc_tail(a, L, tot_s, tot_c, s_max);
    } else {
      a.blk_a[blockIdx.x] = tot_s;
      a.blk_b[blockIdx.x] = tot_c;
      a.blk_c[blockIdx.x] = s_max;
    }
  }
}

// One workgroup: block totals -> exclusive block offsets (in place), then
// alloc_tail.
__global__ __launch_bounds__(kScanThreads) void k_alloc2(LevelArgs a, int nb) {
  __shared__ int s_scan[16];
  const LevelCtl L = *a.ctl;
  if (!L.active) {
    if (threadIdx.x == 0) *a.ctl_next = LevelCtl{L.le, L.le, 0, 0};
    return;
  }
  __shared__ int s_max;
  if (threadIdx.x == 0) s_max = 0;
  __syncthreads();
  int carry_s = 0, carry_c = 0;
  for (int b = 0; b < nb; b += kScanThreads) {
    const int i = b + threadIdx.x;
    if (i < nb) atomicMax(&s_max, a.blk_c[i]);
    const int vs = i < nb ? a.blk_a[i] : 0, vc = i < nb ? a.blk_b[i] : 0;
    int ts, tc;
    const int xs = block_exscan(vs, s_scan, &ts);
    const int xc = block_exscan(vc, s_scan, &tc);
    if (i < nb) {
      a.blk_a[i] = carry_s + xs;
      a.blk_b[i] = carry_c + xc;
    }
    carry_s += ts;
    carry_c += tc;
  }
  if (threadIdx.x == 0) alloc_tail(a, L, carry_s, carry_c, s_max);
}

// Blocks [0, chunk_blocks): one wave per chunk, the stable partition into
// the other point buffer.  The rest: one thread per node, the children.
__global__ void k_split(LevelArgs a, int chunk_blocks) {
  const LevelCtl L = *a.ctl;
  if (!L.active) return;
  if ((int)blockIdx.x >= chunk_blocks) {
    const int j = ((int)blockIdx.x - chunk_blocks) * blockDim.x + threadIdx.x;
    if (j >= L.le - L.lb) return;
    BNode& nd = a.nodes[L.lb + j];
    if (!nd.split) return;
    const int blk = j / kScanThreads;
    const int k = a.blk_a[blk] + nd.ex_split;
    const int li = L.le + 2 * k, ri = li + 1;
    if (ri >= a.cap_nodes) return;  // flagged by k_alloc
    nd.left = li;
    nd.right = ri;
    const int nl = nd.n_left, n = nd.end - nd.start;
    const int c0 = a.blk_b[blk] + nd.ex_chunks;
    a.cbeg_next[2 * k] = c0;
    a.cbeg_next[2 * k + 1] = c0 + (nl + kC - 1) / kC;
    BNode ch;
    memset(&ch, 0, sizeof(ch));
    ch.left = ch.right = -1;
    ch.parent = L.lb + j;
    ch.level = nd.level + 1;
    ch.has_prop = nd.has_prop;
    for (int q = 0; q < 3; ++q) ch.prop[q] = nd.prop[q];
    ch.start = nd.start;
    ch.end = nd.start + nl;
    a.nodes[li] = ch;
    ch.start = nd.start + nl;
    ch.end = nd.start + n;
    a.nodes[ri] = ch;
    return;
  }
  const int c = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = lane_id();
  if (c >= L.n_chunks) return;
  const ChunkRef r = chunk_ref(a, L, c);
  const BNode& nd = a.nodes[L.lb + r.j];
  const bool valid = lane < r.cnt;
  const int i = r.p0 + min(lane, r.cnt - 1);
  const float x = a.px[i], y = a.py[i], z = a.pz[i];
  const double t = a.pt[i];
  int o = i;
  if (nd.split) {
    // left points of the node's earlier chunks: whole 64-chunk groups, then
    // the earlier chunks of this group (integer sums)
    int before_left = 0;
    if (r.m > 1) {
      const int g = (c - r.cb) >> 6;
      for (int b = 0; b < g; b += 64) before_left += wave_sum_int(b + lane < g ? a.gleft[r.cb + b + lane] : 0);
      const int in_group = c - r.cb - kC * g;
      before_left += wave_sum_int(lane < in_group ? a.cleft[r.cb + kC * g + lane] : 0);
    }
    // the next level's chunk-to-node map: the children's cl + cr chunks
    // (m or m + 1 of them) are written by the parent's chunks in order
    if (lane == 0) {
      const int blk = r.j / kScanThreads;
      const int k = a.blk_a[blk] + nd.ex_split;
      const int li = L.le + 2 * k;
      const int c0 = a.blk_b[blk] + nd.ex_chunks;
      const int cl = (nd.n_left + kC - 1) / kC;
      const int cr = (nd.end - nd.start - nd.n_left + kC - 1) / kC;
      const int t = c - r.cb;
      if (t < cl + cr) a.chunk_node_next[c0 + t] = t < cl ? li : li + 1;
      if (t == r.m - 1 && cl + cr == r.m + 1) a.chunk_node_next[c0 + r.m] = li + 1;
    }
    const int before_right = (c - r.cb) * kC - before_left;
    // stable partition: a point's rank among the chunk's points of its side
    const bool left = valid && tg::goes_left(nd.dir, nd.mean, x, y, z);
    const unsigned long long mk = __ballot(left);
    const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
    const int rank_l = __popcll(mk & below);
    o = left ? nd.start + before_left + rank_l : nd.start + nd.n_left + before_right + (lane - rank_l);
  }
  if (valid) {
    a.qx[o] = x;
    a.qy[o] = y;
    a.qz[o] = z;
    a.qt[o] = t;
  }
}

// ---- subtree phase ---------------------------------------------------------------
//
// Once every node of a level holds at most kSubLimit points, one workgroup
// per node finishes its subtree: the points live in LDS, and every depth runs
// the level pipeline's steps with barriers instead of launches -- 64-point
// chunk sums by a wave each (the same pairwise shuffles), the node sums in
// the same pairwise order (at most 32 chunks: one group), the labelling, the
// split counts, children in breadth-first order within the subtree, and the
// stable partition by a prefix sum over the points.  The subtree's nodes are
// staged with their depth and their index within the depth; the merge puts
// depth d of subtree s after depth d of subtrees 0..s-1, which is the
// breadth-first order of the whole tree.

struct SubArgs {
  BNode* nodes;          // roots at [lb, lb + n_roots)
  int lb;
  int level0;            // level of the roots
  BNode* stage;
  int* depth_cnt;
  const float *px, *py, *pz;  // the roots' points
  const double* pt;
  float *ox, *oy, *oz;        // final point order
  double* ot;
  Status* st;
};

constexpr int kSubWaves = kSubThreads / 64;

// One node's sums in the canonical order by one wave, from the points in LDS:
// each 64-point chunk by the pairwise shuffles, the chunk sums (at most 32)
// through the wave's LDS scratch and one more pairwise round.  item(i, v)
// fills the K values of point i.  Every lane ends with the node sums.
// The canonical sum of one chunk (cnt <= 64 values from index base) by one
// row of 16 lanes: lane j of the row takes values 4j .. 4j+3 and forms the
// first two rounds of the pairwise tree in registers ((v0 + v1) + (v2 + v3),
// each addition only when its right operand exists); value strides 4, 8, 16
// and 32 are lane strides 1, 2, 4 and 8 inside the row (DPP).  The row's
// lane 0 ends with the sum.
template <int K, typename Item>
__device__ __forceinline__ void row_chunk_sum(int base, int cnt, int j, const Item& item, double (&a0)[K]) {
  double a1[K], a2[K], a3[K];
  item(base + min(4 * j + 0, cnt - 1), a0);
  item(base + min(4 * j + 1, cnt - 1), a1);
  item(base + min(4 * j + 2, cnt - 1), a2);
  item(base + min(4 * j + 3, cnt - 1), a3);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (4 * j + 1 < cnt) a0[k] += a1[k];
    if (4 * j + 3 < cnt) a2[k] += a3[k];
    if (4 * j + 2 < cnt) a0[k] += a2[k];
  }
  row_round<1>(a0, cnt, j);
  row_round<2>(a0, cnt, j);
  row_round<4>(a0, cnt, j);
  row_round<8>(a0, cnt, j);
}

// Chunks t0 .. t0+3 of a node, one per row; each row's lane 0 stores its
// chunk sum in scratch[t].
template <int K, typename Item>
__device__ __forceinline__ void chunk_sums_rows(int ns, int nn, int t0, int m, double (*scratch)[6],
                                                const Item& item) {
  const int lane = lane_id(), j = lane & 15;
  const int t = t0 + (lane >> 4);
  const int tc = min(t, m - 1);  // rows past the last chunk load a real one
  double a0[K];
  row_chunk_sum<K>(ns + tc * kC, min(kC, nn - tc * kC), j, item, a0);
  if (j == 0 && t < m)
    for (int k = 0; k < K; ++k) scratch[t][k] = a0[k];
}

template <int K, typename Item>
__device__ __forceinline__ void node_sums(int ns, int nn, double (*scratch)[6], const Item& item, double (&out)[K]) {
  const int lane = lane_id();
  const int m = (nn + kC - 1) / kC;
  for (int t0 = 0; t0 < m; t0 += 4) chunk_sums_rows<K>(ns, nn, t0, m, scratch, item);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  double v[K];
  for (int k = 0; k < K; ++k) v[k] = scratch[min(lane, m - 1)][k];
  wave_pairwise<K>(v, m);
  for (int k = 0; k < K; ++k) out[k] = v[k];
  __builtin_amdgcn_wave_barrier();  // the scratch is reused by the next node
}

// LIMIT: the largest subtree this launch takes; it sizes the LDS, so a
// launch for small subtrees keeps several workgroups on a CU.  Subtrees of
// at most `lo` points belong to another launch.
template <int LIMIT>
__global__ __launch_bounds__(kSubThreads) void k_subtree(SubArgs a, kdh_tree_gen_params P, int lo) {
  __shared__ float sx[2][LIMIT], sy[2][LIMIT], sz[2][LIMIT];
  __shared__ double stt[2][LIMIT];
  __shared__ double s_chunk[kSubWaves][LIMIT / kC][6];
  __shared__ int s_scan[16];
  const int tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
  const int sidx = blockIdx.x;
  BNode* root = &a.nodes[a.lb + sidx];
  const int p_begin = root->start, n = root->end - root->start;
  if (n <= lo || n > LIMIT) return;
  BNode* S = a.stage + 2 * (size_t)p_begin;
  int* DC = a.depth_cnt + p_begin + sidx;
  const double ts0 = a.st->ts_start;
  for (int i = tid; i < n; i += kSubThreads) {
    sx[0][i] = a.px[p_begin + i];
    sy[0][i] = a.py[p_begin + i];
    sz[0][i] = a.pz[p_begin + i];
    stt[0][i] = a.pt[p_begin + i];
  }
  if (tid == 0) {
    BNode r = *root;
    r.left = r.right = -1;
    r.ex_split = 0;  // index within its depth
    S[0] = r;
  }
  int q0 = 0, q1 = 1, d = 0;
  __syncthreads();
  for (;;) {
    const int m_d = q1 - q0;
    if (tid == 0) DC[d] = m_d;
    // moments: a node of more than one chunk by a wave, ...
    for (int k = wave; k < m_d; k += kSubWaves) {
      BNode& nd = S[q0 + k];
      const int ns = nd.start - p_begin, nn = nd.end - nd.start;
      if (nn <= kC) continue;
      double s1[5];
      node_sums<5>(ns, nn, s_chunk[wave], [&](int i, double (&v)[5]) {
        const double dd = stt[0][i] - ts0;
        v[0] = (double)sx[0][i];
        v[1] = (double)sy[0][i];
        v[2] = (double)sz[0][i];
        v[3] = dd;
        v[4] = dd * dd;
      }, s1);
      const double inv = 1.0 / (double)nn;
      const double mu0 = s1[0] * inv, mu1 = s1[1] * inv, mu2 = s1[2] * inv;
      double c2[6];
      node_sums<6>(ns, nn, s_chunk[wave], [&](int i, double (&v)[6]) {
        const double d0 = (double)sx[0][i] - mu0, d1 = (double)sy[0][i] - mu1, d2 = (double)sz[0][i] - mu2;
        v[0] = d0 * d0;
        v[1] = d0 * d1;
        v[2] = d0 * d2;
        v[3] = d1 * d1;
        v[4] = d1 * d2;
        v[5] = d2 * d2;
      }, c2);
      if (lane == 0) {
        for (int q = 0; q < 5; ++q) nd.s1[q] = s1[q];
        for (int q = 0; q < 6; ++q) nd.c2[q] = c2[q];
      }
    }
    // ... a single-chunk node (its sum is its chunk's sum) by a row of 16
    // lanes, four nodes to a wave
    {
      const int row = lane >> 4, j = lane & 15;
      for (int k0 = 4 * wave; k0 < m_d; k0 += 4 * kSubWaves) {
        const int k = min(k0 + row, m_d - 1);
        BNode& nd = S[q0 + k];
        const int ns = nd.start - p_begin, nn = nd.end - nd.start;
        const bool mine = k0 + row < m_d && nn <= kC;
        if (__ballot(mine) == 0) continue;
        const int cnt = min(nn, kC);  // a larger node's first chunk: loaded, not stored
        double s1[5];
        row_chunk_sum<5>(ns, cnt, j, [&](int i, double (&v)[5]) {
          const double dd = stt[0][i] - ts0;
          v[0] = (double)sx[0][i];
          v[1] = (double)sy[0][i];
          v[2] = (double)sz[0][i];
          v[3] = dd;
          v[4] = dd * dd;
        }, s1);
        const int src = lane & ~15;
        const double inv = 1.0 / (double)nn;
        const double mu0 = __shfl(s1[0], src, 64) * inv, mu1 = __shfl(s1[1], src, 64) * inv,
                     mu2 = __shfl(s1[2], src, 64) * inv;
        double c2[6];
        row_chunk_sum<6>(ns, cnt, j, [&](int i, double (&v)[6]) {
          const double d0 = (double)sx[0][i] - mu0, d1 = (double)sy[0][i] - mu1, d2 = (double)sz[0][i] - mu2;
          v[0] = d0 * d0;
          v[1] = d0 * d1;
          v[2] = d0 * d2;
          v[3] = d1 * d1;
          v[4] = d1 * d2;
          v[5] = d2 * d2;
        }, c2);
        if (mine && j == 0) {
          for (int q = 0; q < 5; ++q) nd.s1[q] = s1[q];
          for (int q = 0; q < 6; ++q) nd.c2[q] = c2[q];
        }
      }
    }
    __syncthreads();
    // labels, a thread per node
    for (int k = tid; k < m_d; k += kSubThreads) {
      BNode& nd = S[q0 + k];
      tg::Label lab;
      tg::label_node(P, (long)(nd.end - nd.start), nd.s1, nd.c2, nd.level == 0, nd.prop, nd.has_prop, lab);
      for (int q = 0; q < 3; ++q) {
        nd.mean[q] = lab.mean[q];
        nd.dir[q] = lab.dir[q];
        nd.evec0[q] = lab.evec0[q];
        nd.prop[q] = lab.prop[q];
      }
      nd.has_prop = lab.has_prop;
      nd.is_node = lab.is_node;
      nd.dts_mean = lab.dts_mean;
      nd.dts_cov = lab.dts_cov;
    }
    __syncthreads();
    // a wave per node: left count, split decision, stable partition into the
    // second buffer and back
    for (int k = wave; k < m_d; k += kSubWaves) {
      BNode& nd = S[q0 + k];
      const int ns = nd.start - p_begin, nn = nd.end - nd.start;
      const int m = (nn + kC - 1) / kC;
      int nl = 0;
      if (nd.is_node) {
        for (int t = 0; t < m; ++t) {
          const int i = ns + t * kC + lane;
          const bool left = (t * kC + lane < nn) && tg::goes_left(nd.dir, nd.mean, sx[0][i], sy[0][i], sz[0][i]);
          nl += __popcll(__ballot(left));
        }
      }
      const int split = (nd.is_node && nl > 0 && nl < nn) ? 1 : 0;
      if (lane == 0) {
        if (nd.is_node && !split)  // one-sided split: a leaf with its normal
          for (int q = 0; q < 3; ++q) nd.dir[q] = nd.evec0[q];
        nd.split = split;
        nd.n_left = nd.is_node ? nl : 0;
      }
      if (!split) continue;
      int lbefore = 0;
      for (int t = 0; t < m; ++t) {
        const int j = t * kC + lane, i = ns + j;
        const bool valid = j < nn;
        const float x = valid ? sx[0][i] : 0.0f, y = valid ? sy[0][i] : 0.0f, z = valid ? sz[0][i] : 0.0f;
        const bool left = valid && tg::goes_left(nd.dir, nd.mean, x, y, z);
        const unsigned long long mk = __ballot(left);
        const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
        const int rl = lbefore + __popcll(mk & below);
        if (valid) {
          const int o = ns + (left ? rl : nl + (j - rl));
          sx[1][o] = x;
          sy[1][o] = y;
          sz[1][o] = z;
          stt[1][o] = stt[0][i];
        }
        lbefore += __popcll(mk);
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      for (int j = lane; j < nn; j += 64) {
        sx[0][ns + j] = sx[1][ns + j];
        sy[0][ns + j] = sy[1][ns + j];
        sz[0][ns + j] = sz[1][ns + j];
        stt[0][ns + j] = stt[1][ns + j];
      }
    }
    __syncthreads();
    // children in order, a thread per node
    int carry = 0;
    for (int b = 0; b < m_d; b += kSubThreads) {
      const int k = b + tid;
      const int split = k < m_d ? S[q0 + k].split : 0;
      int tot;
      const int ex = block_exscan(split, s_scan, &tot);
      if (split) {
        BNode& nd = S[q0 + k];
        const int kk = carry + ex;
        const int li = q1 + 2 * kk;
        nd.left = li;
        nd.right = li + 1;
        for (int side = 0; side < 2; ++side) {
          BNode& ch = S[li + side];
          memset(&ch, 0, sizeof(BNode));
          ch.left = ch.right = -1;
          ch.parent = q0 + k;
          ch.level = nd.level + 1;
          ch.has_prop = nd.has_prop;
          for (int q = 0; q < 3; ++q) ch.prop[q] = nd.prop[q];
          ch.start = side ? nd.start + nd.n_left : nd.start;
          ch.end = side ? nd.end : nd.start + nd.n_left;
          ch.ex_split = 2 * kk + side;
        }
      }
      carry += tot;
    }
    const int n_split = carry;
    __syncthreads();
    if (n_split == 0) break;
    q0 = q1;
    q1 += 2 * n_split;
    ++d;
  }
  for (int i = tid; i < n; i += kSubThreads) {
    a.ox[p_begin + i] = sx[0][i];
    a.oy[p_begin + i] = sy[0][i];
    a.oz[p_begin + i] = sz[0][i];
    a.ot[p_begin + i] = stt[0][i];
  }
  if (tid == 0) {
    S[0].n_depths = d + 1;
    S[0].n_staged = q1;
    atomicMax(&a.st->max_depth, d + 1);
  }
}

// One workgroup: per depth, the prefix sum over the subtrees of their node
// counts (in place: each count becomes the subtree's offset in the level),
// and the level starts.
__global__ __launch_bounds__(kScanThreads) void k_sub_levels(SubArgs a, int n_roots, int* level_begin) {
  __shared__ int s_scan[16];
  const int D = a.st->max_depth;
  for (int d = 0; d < D; ++d) {
    int carry = 0;
    for (int b = 0; b < n_roots; b += kScanThreads) {
      const int s = b + threadIdx.x;
      int v = 0;
      int* slot = nullptr;
      if (s < n_roots) {
        const int p_begin = a.nodes[a.lb + s].start;
        const BNode& r = a.stage[2 * (size_t)p_begin];
        if (d < r.n_depths) {
          slot = a.depth_cnt + p_begin + s + d;
          v = *slot;
        }
      }
      int tot;
      const int ex = block_exscan(v, s_scan, &tot);
      if (slot) *slot = carry + ex;
      carry += tot;
    }
    if (threadIdx.x == 0) level_begin[a.level0 + d + 1] = level_begin[a.level0 + d] + carry;
    __syncthreads();
  }
  if (threadIdx.x == 0) a.st->n_built = level_begin[a.level0 + D];
}

// One workgroup per subtree: global indices of its staged nodes, then the
// nodes with global links into the level-ordered array.
__global__ void k_sub_place(SubArgs a, const int* level_begin) {
  const int sidx = blockIdx.x;
  const BNode* root = &a.nodes[a.lb + sidx];
  const int p_begin = root->start;
  const int parent_of_root = root->parent;
  BNode* S = a.stage + 2 * (size_t)p_begin;
  const int* DC = a.depth_cnt + p_begin + sidx;
  const int n_staged = S[0].n_staged;
  __syncthreads();
  for (int i = threadIdx.x; i < n_staged; i += blockDim.x) {
    const int dd = S[i].level - a.level0;
    S[i].gidx = level_begin[a.level0 + dd] + DC[dd] + S[i].ex_split;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n_staged; i += blockDim.x) {
    const BNode& r = S[i];
    BNode& g = a.nodes[r.gidx];
    g = r;
    g.left = r.left >= 0 ? S[r.left].gidx : -1;
    g.right = r.right >= 0 ? S[r.right].gidx : -1;
    g.parent = i == 0 ? parent_of_root : S[r.parent].gidx;
  }
}

// ---- prune and breadth-first numbering -------------------------------------------

// TreeCPU_::_prune bottom-up: every leaf decides its keep flag and climbs.
// A parent's word collects 2 + keep per child; the second child to arrive
// reads its sibling's flag from the atomic's return value and decides the
// parent.
__global__ void k_keep(BNode* nodes, int* keep_word, int n_built, const Status* st) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (st->overflow || i >= n_built) return;
  BNode* nd = &nodes[i];
  if (nd->split) return;
  int keep = tg::good_leaf(nd->mean, nd->dir) ? 1 : 0;
  nd->keep = keep;
  int p = nd->parent;
  while (p >= 0) {
    const int old = atomicAdd(&keep_word[p], 2 + keep);
    if (old == 0) return;  // the sibling decides
    keep_word[p] = 0;
    keep = ((old - 2) | keep) & 1;
    nodes[p].keep = keep;
    p = nodes[p].parent;
  }
}

// Prefix sums of the keep and kept-leaf flags over the level order within
// each 1024-node block.
__global__ __launch_bounds__(kScanThreads) void k_number(BNode* nodes, int* blk_a, int* blk_b, int n_built,
                                                         const Status* st) {
  __shared__ int s_scan[16];
  if (st->overflow) return;
  const int i = blockIdx.x * kScanThreads + threadIdx.x;
  int keep = 0, leaf = 0;
  if (i < n_built) {
    keep = nodes[i].keep;
    leaf = keep && !nodes[i].split;
  }
  int tk, tl;
  const int ek = block_exscan(keep, s_scan, &tk);
  const int el = block_exscan(leaf, s_scan, &tl);
  if (i < n_built) {
    nodes[i].ex_keep = ek;
    nodes[i].ex_leaf = el;
  }
  if (threadIdx.x == 0) {
    blk_a[blockIdx.x] = tk;
    blk_b[blockIdx.x] = tl;
  }
}

// One workgroup: the block offsets and the totals.
__global__ __launch_bounds__(kScanThreads) void k_number2(const BNode* nodes, int* blk_a, int* blk_b, int nb,
                                                          Status* st) {
  __shared__ int s_scan[16];
  if (st->overflow) return;
  int ck = 0, cl = 0;
  for (int b = 0; b < nb; b += kScanThreads) {
    const int j = b + threadIdx.x;
    const int vk = j < nb ? blk_a[j] : 0, vl = j < nb ? blk_b[j] : 0;
    int sk, sl;
    const int xk = block_exscan(vk, s_scan, &sk);
    const int xl = block_exscan(vl, s_scan, &sl);
    if (j < nb) {
      blk_a[j] = ck + xk;
      blk_b[j] = cl + xl;
    }
    ck += sk;
    cl += sl;
  }
  if (threadIdx.x == 0) {
    st->n_out = ck;
    st->n_leaves_out = cl;
    st->root_kept = nodes[0].keep;
  }
}

struct EmitArgs {
  const BNode* nodes;
  const int* blk_a;    // kept-node offsets per prefix-sum block
  const int* blk_b;    // kept-leaf offsets per prefix-sum block
  int n_nodes;
  float* dnodes;       // 8 floats per node
  float* dleaves;      // 8 floats per leaf
  float* ddts;         // per node
  kdh_node3f* full;    // Node_ records
  uint64_t* leaf_ids;  // leaf list
};

__device__ __forceinline__ int new_index(const EmitArgs& a, int i) {
  return a.blk_a[i / kScanThreads] + a.nodes[i].ex_keep;
}

__global__ void k_emit(EmitArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n_nodes) return;
  const BNode& nd = a.nodes[i];
  if (!nd.keep) return;
  const int o = new_index(a, i);
  int l = -1, r = -1;
  if (nd.split) {
    if (a.nodes[nd.left].keep) l = new_index(a, nd.left);
    if (a.nodes[nd.right].keep) r = new_index(a, nd.right);
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
  if (!nd.split) {
    const int li = a.blk_b[i / kScanThreads] + nd.ex_leaf;
    float* lr = a.dleaves + (size_t)li * 8;
    lr[0] = nd.mean[0];
    lr[1] = nd.mean[1];
    lr[2] = nd.mean[2];
    lr[3] = nd.dir[0];
    lr[4] = nd.dir[1];
    lr[5] = nd.dir[2];
    lr[6] = (float)nd.dts_mean;
    lr[7] = __int_as_float(o);
    a.leaf_ids[li] = (uint64_t)o;
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
  const int n_blk = (int)((2 * N + kScanThreads - 1) / kScanThreads) + 1;
  const size_t zero_bytes = a256(2 * N * 8) + a256(N * 8) + a256(2 * N * 4);  // arrival words
  const size_t bytes = zero_bytes + 3 * 2 * a256(N * 4) + 2 * a256(N * 8) + a256(2 * N * sizeof(BNode)) +
                       2 * a256(N * 6 * 8) + 2 * a256(N * 4) + 2 * a256(N * 4) + 3 * a256((size_t)n_blk * 4) +
                       a256(2 * N * sizeof(BNode)) + a256(2 * N * 4) +
                       2 * a256((N + 1) * 4) + a256((kMaxLevels + 2) * 4) + a256(2 * sizeof(LevelCtl)) +
                       a256(sizeof(Status)) + a256(3 * N * 4) + a256(N * 8) + a256(8);
  void* mem = nullptr;
  hipError_t e = hipMalloc(&mem, bytes);
  if (e != hipSuccess)
    return kdh_fail(ctx, KDH_ERR_NOMEM, std::string("kdh_tree_build: workspace: ") + hipGetErrorString(e));
  // arrival counters start at zero and every use leaves them at zero
  e = hipMemset(mem, 0, zero_bytes);
  if (e != hipSuccess) {
    hipFree(mem);
    return kdh_hip_fail(ctx, e, "kdh_tree_build: workspace clear");
  }
  ws = new Ws();
  ws->cap = cap;
  ws->mem = mem;
  ws->n_blk = n_blk;
  char* p = (char*)mem;
  auto take = [&](size_t b) {
    char* q = p;
    p += a256(b);
    return q;
  };
  ws->cnt64 = (unsigned long long*)take(2 * N * 8);
  ws->cnt64_grp = (unsigned long long*)take(N * 8);
  ws->keep_word = (int*)take(2 * N * 4);
  for (int k = 0; k < 2; ++k) {
    ws->px[k] = (float*)take(N * 4);
    ws->py[k] = (float*)take(N * 4);
    ws->pz[k] = (float*)take(N * 4);
    ws->pt[k] = (double*)take(N * 8);
  }
  ws->nodes = (BNode*)take(2 * N * sizeof(BNode));
  ws->csum = (double*)take(N * 6 * 8);
  ws->gsum = (double*)take(N * 6 * 8);
  ws->cleft = (int*)take(N * 4);
  ws->gleft = (int*)take(N * 4);
  ws->chunk_node[0] = (int*)take(N * 4);
  ws->chunk_node[1] = (int*)take(N * 4);
  ws->blk_a = (int*)take((size_t)n_blk * 4);
  ws->blk_b = (int*)take((size_t)n_blk * 4);
  ws->blk_c = (int*)take((size_t)n_blk * 4);
  ws->stage = (BNode*)take(2 * N * sizeof(BNode));
  ws->depth_cnt = (int*)take(2 * N * 4);
  ws->cbeg[0] = (int*)take((N + 1) * 4);
  ws->cbeg[1] = (int*)take((N + 1) * 4);
  ws->level_begin = (int*)take((kMaxLevels + 2) * 4);
  ws->ctl = (LevelCtl*)take(2 * sizeof(LevelCtl));
  ws->status = (Status*)take(sizeof(Status));
  ws->stage_xyz = (float*)take(3 * N * 4);
  ws->stage_t = (double*)take(N * 8);
  ws->ts_key = (unsigned long long*)take(8);
  e = hipMemset(ws->ts_key, 0xFF, 8);
  if (e != hipSuccess) {
    kdh_tree_build_ws_free(ctx);
    return kdh_hip_fail(ctx, e, "kdh_tree_build: workspace clear");
  }
  e = hipHostMalloc((void**)&ws->h_status, sizeof(Status) * (kMaxLevels / 2 + 2) + 64, hipHostMallocDefault);
  if (e != hipSuccess) {
    kdh_tree_build_ws_free(ctx);
    return kdh_fail(ctx, KDH_ERR_NOMEM, "kdh_tree_build: pinned status");
  }
  ws->h_count = reinterpret_cast<int*>(ws->h_status + (kMaxLevels / 2 + 2));
  return KDH_OK;
}

}  // namespace

void kdh_tree_build_ws_free(kdh_ctx* ctx) {
  Ws*& ws = reinterpret_cast<Ws*&>(ctx->tree_build_ws);
  if (!ws) return;
  if (ws->first_exec) hipGraphExecDestroy(ws->first_exec);
  for (int k = 0; k < 2; ++k) {
    if (ws->aux[k]) hipStreamDestroy(ws->aux[k]);
    if (ws->join[k]) hipEventDestroy(ws->join[k]);
  }
  if (ws->fork) hipEventDestroy(ws->fork);
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
  if (P.label_mode != 0)
    return kdh_fail(ctx, KDH_ERR_INVALID,
                    "kdh_tree_build: label_mode 1 (the reference's own arithmetic) is built on the host "
                    "(kdh_host_tree_build)");
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
                     ws->pt[0], ws->ts_key);
  hipLaunchKernelGGL(k_init, dim3(1), dim3(kScanThreads), 0, s, ws->ts_key, n, ws->nodes, ws->ctl, ws->cbeg[0],
                     ws->chunk_node[0], ws->level_begin, ws->status);
  KDH_HIP_TRY(ctx, hipGetLastError());

  // Launch shapes come from the workspace capacity, not from n, so that a
  // batch of levels can be replayed as a graph for any cloud it holds.
  const int64_t cap = ws->cap;
  const int64_t cap_chunks = (cap + kC - 1) / kC;
  // Levels whose nodes average four or more 64-chunk groups (about 16k
  // points; the first batch reaches nodes of 2048) run the first round of
  // their reductions across many waves (k_gsum).
  int first_batch = kBatchLevels;
  while (first_batch < 64 && ((int64_t)kSubLimit << (first_batch - 2)) < (int64_t)n) ++first_batch;
  const int grouped_levels = first_batch - 4;  // levels [0, grouped_levels)
  auto enqueue_level = [&](int L) -> hipError_t {
    const int grouped = L < grouped_levels ? 1 : 0;
    const int cur = L & 1, nxt = cur ^ 1;
    const int64_t width_bound = std::min<int64_t>((int64_t)1 << std::min(L, 30), cap);
    const int64_t chunk_bound = std::min<int64_t>(cap, cap_chunks + width_bound);
    const int gc = (int)((chunk_bound + 3) / 4);  // one wave per chunk, four per block
    const int gn = (int)((width_bound + 255) / 256);
    const int ga = (int)((width_bound + kScanThreads - 1) / kScanThreads);
    const int gw = (int)std::min<int64_t>((width_bound + 3) / 4, 4096);  // one wave per node
    LevelArgs a;
    a.nodes = ws->nodes;
    a.ctl = ws->ctl + cur;
    a.ctl_next = ws->ctl + nxt;
    a.cbeg = ws->cbeg[cur];
    a.cbeg_next = ws->cbeg[nxt];
    a.px = ws->px[cur];
    a.py = ws->py[cur];
    a.pz = ws->pz[cur];
    a.pt = ws->pt[cur];
    a.qx = ws->px[nxt];
    a.qy = ws->py[nxt];
    a.qz = ws->pz[nxt];
    a.qt = ws->pt[nxt];
    a.chunk_node = ws->chunk_node[cur];
    a.chunk_node_next = ws->chunk_node[nxt];
    a.csum = ws->csum;
    a.gsum = ws->gsum;
    a.cleft = ws->cleft;
    a.gleft = ws->gleft;
    a.cnt64 = ws->cnt64;
    a.cnt64_grp = ws->cnt64_grp;
    a.blk_a = ws->blk_a;
    a.blk_b = ws->blk_b;
    a.blk_c = ws->blk_c;
    a.level_begin = ws->level_begin;
    a.st = ws->status;
    a.level = L;
    a.cap_nodes = cap_nodes;
    const dim3 g_grp((unsigned)width_bound, 8);  // a node's groups over 8 x 16 waves
    hipLaunchKernelGGL(k_mom1, dim3(gc), dim3(256), 0, s, a);
    if (grouped) hipLaunchKernelGGL(k_gsum<5>, g_grp, dim3(1024), 0, s, a);
    hipLaunchKernelGGL(k_red1, dim3(gw), dim3(256), 0, s, a, grouped);
    hipLaunchKernelGGL(k_mom2, dim3(gc), dim3(256), 0, s, a);
    if (grouped) hipLaunchKernelGGL(k_gsum<6>, g_grp, dim3(1024), 0, s, a);
    hipLaunchKernelGGL(k_red2, dim3(gw), dim3(256), 0, s, a, P, grouped);
    hipLaunchKernelGGL(k_count, dim3(gc), dim3(256), 0, s, a);
    if (ga == 1) {
      hipLaunchKernelGGL(k_alloc<true>, dim3(1), dim3(kScanThreads), 0, s, a);
    } else {
      hipLaunchKernelGGL(k_alloc<false>, dim3(ga), dim3(kScanThreads), 0, s, a);
      hipLaunchKernelGGL(k_alloc2, dim3(1), dim3(kScanThreads), 0, s, a, ga);
    }
    hipLaunchKernelGGL(k_split, dim3(gc + gn), dim3(256), 0, s, a, gc);
    return hipGetLastError();
  };

  // Levels are queued one batch ahead of the status the host has seen; the
  // levels queued past the last one find no nodes and return at once.  The
  // first batch reaches the level where balanced splits hand over to the
  // subtree phase; later batches are short.
  int queued = 0, n_levels = -1;
  Status seen;
  memset(&seen, 0, sizeof(seen));
  if (n <= kSubLimit) {  // the root is a subtree already
    n_levels = 0;
    seen.handoff = 1;
    seen.handoff_level = 0;
    seen.handoff_lb = 0;
    seen.handoff_n = 1;
  }
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
    if (evs.empty()) {
      // the first batch: a cached graph of its launches
      if (!ws->first_exec || ws->first_levels != first_batch || memcmp(&ws->first_params, &P, sizeof(P)) != 0) {
        if (ws->first_exec) hipGraphExecDestroy(ws->first_exec);
        ws->first_exec = nullptr;
        hipGraph_t graph = nullptr;
        e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
        for (int k = 0; k < first_batch && e == hipSuccess; ++k) e = enqueue_level(k);
        const hipError_t ce = hipStreamEndCapture(s, &graph);
        if (e == hipSuccess) e = ce;
        if (e == hipSuccess) e = hipGraphInstantiate(&ws->first_exec, graph, nullptr, nullptr, 0);
        if (graph) hipGraphDestroy(graph);
        if (e == hipSuccess) {
          ws->first_levels = first_batch;
          ws->first_params = P;
        } else {
          ws->first_exec = nullptr;
        }
      }
      if (e == hipSuccess) e = hipGraphLaunch(ws->first_exec, s);
      queued = first_batch;
    } else {
      for (int k = 0; k < 2 && e == hipSuccess; ++k) e = enqueue_level(queued++);
    }
    hipEvent_t ev = nullptr;
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    if (e == hipSuccess) evs.push_back(ev);
    const int b = (int)evs.size() - 1;
    if (e == hipSuccess)
      e = hipMemcpyAsync(ws->h_status + b, ws->status, sizeof(Status), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipEventRecord(ev, s);
    if (e == hipSuccess && b >= 1) {
      e = hipEventSynchronize(evs[b - 1]);
      if (e == hipSuccess && ws->h_status[b - 1].last_level >= 0) {
        n_levels = ws->h_status[b - 1].last_level;
        seen = ws->h_status[b - 1];
      }
    }
    if (e != hipSuccess) {
      release();
      return kdh_hip_fail(ctx, e, "kdh_tree_build: level queue");
    }
  }
  release();

  int n_built = 0;
  if (seen.handoff) {
    // the subtree phase: the roots' levels on, then the merge into level order
    SubArgs sa;
    sa.nodes = ws->nodes;
    sa.lb = seen.handoff_lb;
    sa.level0 = seen.handoff_level;
    sa.stage = ws->stage;
    sa.depth_cnt = ws->depth_cnt;
    const int cur = seen.handoff_level & 1;
    sa.px = ws->px[cur];
    sa.py = ws->py[cur];
    sa.pz = ws->pz[cur];
    sa.pt = ws->pt[cur];
    sa.ox = ws->px[0];
    sa.oy = ws->py[0];
    sa.oz = ws->pz[0];
    sa.ot = ws->pt[0];
    sa.st = ws->status;
    if (seen.handoff_level + 2 + kSubLimit > kMaxLevels + 2)
      return kdh_fail(ctx, KDH_ERR_INVALID, "kdh_tree_build: deeper than 4096 levels");
    // Subtrees by size class, each launch with LDS for its largest: one
    // round of workgroups holds every subtree when there are few; else the
    // small ones (the most) run several to a CU, and the three classes run
    // side by side on three streams.
    if (seen.handoff_n <= 256) {
      hipLaunchKernelGGL(k_subtree<kSubLimit>, dim3(seen.handoff_n), dim3(kSubThreads), 0, s, sa, P, 0);
    } else {
      for (int k = 0; k < 2; ++k) {
        if (!ws->aux[k]) KDH_HIP_TRY(ctx, hipStreamCreateWithFlags(&ws->aux[k], hipStreamNonBlocking));
        if (!ws->join[k]) KDH_HIP_TRY(ctx, hipEventCreateWithFlags(&ws->join[k], hipEventDisableTiming));
      }
      if (!ws->fork) KDH_HIP_TRY(ctx, hipEventCreateWithFlags(&ws->fork, hipEventDisableTiming));
      KDH_HIP_TRY(ctx, hipEventRecord(ws->fork, s));
      for (int k = 0; k < 2; ++k) KDH_HIP_TRY(ctx, hipStreamWaitEvent(ws->aux[k], ws->fork, 0));
      hipLaunchKernelGGL(k_subtree<kSubLimit>, dim3(seen.handoff_n), dim3(kSubThreads), 0, s, sa, P, 1024);
      hipLaunchKernelGGL(k_subtree<1024>, dim3(seen.handoff_n), dim3(kSubThreads), 0, ws->aux[0], sa, P, 512);
      hipLaunchKernelGGL(k_subtree<512>, dim3(seen.handoff_n), dim3(kSubThreads), 0, ws->aux[1], sa, P, 0);
      for (int k = 0; k < 2; ++k) {
        KDH_HIP_TRY(ctx, hipEventRecord(ws->join[k], ws->aux[k]));
        KDH_HIP_TRY(ctx, hipStreamWaitEvent(s, ws->join[k], 0));
      }
    }
    hipLaunchKernelGGL(k_sub_levels, dim3(1), dim3(kScanThreads), 0, s, sa, seen.handoff_n, ws->level_begin);
    hipLaunchKernelGGL(k_sub_place, dim3(seen.handoff_n), dim3(kSubThreads), 0, s, sa, ws->level_begin);
    KDH_HIP_TRY(ctx, hipGetLastError());
    KDH_HIP_TRY(ctx, hipMemcpyAsync(ws->h_status, ws->status, sizeof(Status), hipMemcpyDeviceToHost, s));
    KDH_HIP_TRY(ctx, hipStreamSynchronize(s));
    n_built = ws->h_status[0].n_built;
  } else {
    // nodes of the unpruned tree: level_begin[n_levels]
    KDH_HIP_TRY(ctx, hipMemcpyAsync(ws->h_count, ws->level_begin + n_levels, sizeof(int), hipMemcpyDeviceToHost, s));
    KDH_HIP_TRY(ctx, hipStreamSynchronize(s));
    n_built = *ws->h_count;
  }
  if (n_built < 1 || n_built > cap_nodes) return kdh_fail(ctx, KDH_ERR_INVALID, "kdh_tree_build: node capacity exceeded");
  const int gp = (n_built + kScanThreads - 1) / kScanThreads;
  if (gp > ws->n_blk) return kdh_fail(ctx, KDH_ERR_INVALID, "kdh_tree_build: prefix-sum blocks exceeded");
  hipLaunchKernelGGL(k_keep, dim3((n_built + 255) / 256), dim3(256), 0, s, ws->nodes, ws->keep_word, n_built,
                     ws->status);
  hipLaunchKernelGGL(k_number, dim3(gp), dim3(kScanThreads), 0, s, ws->nodes, ws->blk_a, ws->blk_b, n_built,
                     ws->status);
  hipLaunchKernelGGL(k_number2, dim3(1), dim3(kScanThreads), 0, s, ws->nodes, ws->blk_a, ws->blk_b, gp, ws->status);
  KDH_HIP_TRY(ctx, hipGetLastError());
  KDH_HIP_TRY(ctx, hipMemcpyAsync(ws->h_status, ws->status, sizeof(Status), hipMemcpyDeviceToHost, s));
  KDH_HIP_TRY(ctx, hipStreamSynchronize(s));
  const Status st = ws->h_status[0];
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
  if (e != hipSuccess)
    return kdh_fail(ctx, KDH_ERR_NOMEM, std::string("kdh_tree_build: hipMalloc: ") + hipGetErrorString(e));
  char* p = (char*)d;
  EmitArgs ea;
  ea.nodes = ws->nodes;
  ea.blk_a = ws->blk_a;
  ea.blk_b = ws->blk_b;
  ea.n_nodes = n_out > 0 ? n_built : 0;
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
