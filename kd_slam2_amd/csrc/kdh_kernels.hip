// HIP kernels of the scan-alignment hot path (gfx950).
//
// One kernel, `tiles_kernel`, serves every execution form:
//  * the plain per-call build (one alignment, mode 1: write H/b/stats),
//  * the device-resident Gauss-Newton loop (one alignment, mode 0),
//  * the batched / served form (many alignments in one launch, mode 0).
// A launch walks a list of tiles; a tile is 256 consecutive moving leaves of
// one alignment ("pair").  Per tile the workgroup
//   1. descends the fixed tree once per leaf (one lane per leaf) with the top
//      of the tree staged in LDS,
//   2. evaluates the point-to-plane term (kdh_term.h),
//   3. reduces the 34 (ICP) or 97 (CT) accumulator slots across the 64 lanes
//      of each wave with a butterfly reduce-scatter (63 or 127 shuffles per
//      wave instead of slots x 6), then across the 4 waves through LDS,
//   4. writes one FP32 partial row per tile to a fixed slot.
// The last workgroup to finish a pair's tiles (atomic counter) sums that
// pair's partial rows in FP64 in tile order -- the result does not depend on
// which workgroup ran which tile, nor on what else shares the launch -- and
// then either publishes H/b (build mode) or runs the FP64 damped LDLT solve,
// the SE(3) update and the next iteration's FP32 cache (solve mode).
//
// Reference counterparts: ICP_kernel / CTICP_kernel + Reduce_ + host
// oneRound (kd_slam2/src/lib/kd_slam/icp/icp_cuda_impl_.cuh:69-132,
// cticp/cticp_cuda_impl_.cuh:140-223, cuda/reduce_impl_.cuh:10-78,
// icp/icp_impl_.h:53-90, cticp/cticp_impl_.h:48-92).
#include <hip/hip_runtime.h>

#include "kdh_internal.h"
#include "kdh_kernels.h"
#include "kdh_solve.h"

#pragma clang fp contract(off)

namespace kdh {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;

template <bool CT> struct Slots;
template <> struct Slots<false> { static constexpr int N = kIcpSlots; static constexpr int P = 64; };
template <> struct Slots<true> { static constexpr int N = kCtSlots; static constexpr int P = 128; };

// Butterfly reduce-scatter over one wave: on entry every lane holds P values
// (P a multiple of 64); on exit lane l holds the wave sums of slots
// j + (P/64)*l, j < P/64, in v[0..P/64).  Fixed pairing => deterministic.
template <int P>
__device__ __forceinline__ void wave_reduce_scatter(float* v, int lane) {
#pragma unroll
  for (int off = 32, h = P / 2; off >= 1; off >>= 1, h >>= 1) {
    const bool upper = (lane & off) != 0;
#pragma unroll
    for (int i = 0; i < h; ++i) {
      float send = upper ? v[i] : v[i + h];
      float keep = upper ? v[i + h] : v[i];
      v[i] = keep + __shfl_xor(send, off, 64);
    }
  }
}

struct GlobalFetch {
  const float4* __restrict__ nodes;
  __device__ __forceinline__ void operator()(int idx, float4& a, float4& b) const {
    a = nodes[2 * idx];
    b = nodes[2 * idx + 1];
  }
};

struct LdsFetch {
  const float4* __restrict__ nodes;
  const float4* lds;
  int n_lds;
  __device__ __forceinline__ void operator()(int idx, float4& a, float4& b) const {
    if (idx < n_lds) {
      a = lds[2 * idx];
      b = lds[2 * idx + 1];
    } else {
      a = nodes[2 * idx];
      b = nodes[2 * idx + 1];
    }
  }
};

// ---- per-pair final stage (runs on the last workgroup of a pair) ----------

template <bool CT>
__device__ void finish_pair(PairState& ps, const float* __restrict__ partials, double* red) {
  constexpr int N = Slots<CT>::N;
  const int tid = threadIdx.x;
  // FP64 sum of the pair's partial rows, in tile order.
  for (int s = tid; s < N; s += kThreads) {
    double acc = 0.0;
    const float* p = partials + (size_t)ps.tile_first * N + s;
    for (int t = 0; t < ps.n_tiles; ++t) acc += (double)p[(size_t)t * N];
    red[s] = acc;
  }
  __syncthreads();
  if (tid != 0) return;

  constexpr int D = CT ? 12 : 6;
  constexpr int CS = CT ? 90 : 27;  // chi2 slot
  DevStats st;
  st.num_inliers = (int)red[CS + 2 + kInlier];
  st.num_outliers = (int)red[CS + 2 + kOutlier];
  st.num_bad = (int)red[CS + 2 + kBad];
  st.num_err_normal = (int)red[CS + 2 + kErrNormal];
  st.num_err_fn = (int)red[CS + 2 + kErrFn];
  st.pad = 0;
  st.chi2 = red[CS];
  st.chi2_kernelized = red[CS + 1];
  st.pert_pose_norm = 0.0;
  st.pert_vel_norm = 0.0;

  double H[D * D];
  double b[D];
  for (int i = 0; i < D * D; ++i) H[i] = 0.0;
  unpack_diag_d(red, H, D, 0, b);
  if (CT) {
    unpack_diag_d(red + 27, H, D, 6, b);
    for (int r = 0; r < 6; ++r)
      for (int c = 0; c < 6; ++c) {
        H[r * D + 6 + c] = red[54 + r * 6 + c];
        H[(6 + c) * D + r] = red[54 + r * 6 + c];
      }
  }
  const V3d cc = v3((double)ps.cache.c.x, (double)ps.cache.c.y, (double)ps.cache.c.z);
  uncenter_d<D>(H, b, cc);

  const int it = ps.iter;
  if (ps.mode == 1) {
    for (int i = 0; i < D * D; ++i) ps.H[i] = H[i];
    for (int i = 0; i < D; ++i) ps.b[i] = b[i];
    ps.stats_log[0] = st;
    ps.tiles_done = 0;
    return;
  }

  IsoD X = iso_from_34<double>(ps.X);
  // pose prior (tree_aligner_base_.h:139-152)
  if (ps.has_prior != 0.0) {
    IsoD ZXi = iso_from_34<double>(ps.prior_ZXi);
    double e[6], J[36];
    logmap_d(compose(ZXi, X), e);
    dPoseToPoseError_d(ZXi, X, J);
    double OJ[36], Oe[6];
    for (int r = 0; r < 6; ++r) {
      double se = 0.0;
      for (int k = 0; k < 6; ++k) se += ps.prior_info[r * 6 + k] * e[k];
      Oe[r] = se;
      for (int c = 0; c < 6; ++c) {
        double s = 0.0;
        for (int k = 0; k < 6; ++k) s += ps.prior_info[r * 6 + k] * J[k * 6 + c];
        OJ[r * 6 + c] = s;
      }
    }
    for (int r = 0; r < 6; ++r) {
      double sb = 0.0;
      for (int k = 0; k < 6; ++k) sb += J[k * 6 + r] * Oe[k];
      b[r] += sb;
      for (int c = 0; c < 6; ++c) {
        double s = 0.0;
        for (int k = 0; k < 6; ++k) s += J[k * 6 + r] * OJ[k * 6 + c];
        H[r * D + c] += s;
      }
    }
  }
  for (int i = 0; i < D * D; ++i) ps.H[i] = H[i];
  for (int i = 0; i < D; ++i) ps.b[i] = b[i];

  const int good = st.num_inliers + st.num_outliers;
  if (good < ps.sp.min_good_measurements) {
    ps.stats_log[it] = st;
    ps.iter = it + 1;
    ps.status = 2;  // failed: compute() returns -it
    ps.tiles_done = 0;
    return;
  }
  double A[D * D], rhs[D], dx[D];
  for (int i = 0; i < D * D; ++i) A[i] = H[i];
  for (int i = 0; i < D; ++i) {
    A[i * D + i] += (i < 6) ? ps.sp.damping : ps.sp.velocity_damping;
    rhs[i] = -b[i];
  }
  ldlt_solve_d<D>(A, rhs, dx);
  double pn = 0.0, vn = 0.0;
  for (int i = 0; i < 6; ++i) pn += dx[i] * dx[i];
  if (CT)
    for (int i = 6; i < D; ++i) vn += dx[i] * dx[i];
  st.pert_pose_norm = sqrt(pn);
  st.pert_vel_norm = sqrt(vn);
  ps.stats_log[it] = st;
  ps.iter = it + 1;
  if (!ps.tp.stats_mode) {
    IsoD Xn = compose(expmap_d(dx), X);
    iso_to_34(Xn, ps.X);
    if (CT)
      for (int i = 0; i < 6; ++i) ps.v[i] += dx[6 + i];
  }
  // termination: stats mode runs one round; the score criterion
  // (icp_stats_.h:85-102) may stop early; otherwise max_iterations.
  bool stop = ps.tp.stats_mode != 0 || ps.iter >= ps.sp.max_iterations;
  if (!stop && ps.sp.use_term_crit) {
    const float ni = (float)st.num_inliers;
    const int nl = st.num_inliers + st.num_outliers + st.num_bad + st.num_err_normal + st.num_err_fn;
    const float ratio = nl ? ni / (float)nl : 0.0f;
    const float score = ratio / ((float)st.chi2_kernelized / ni + 1e-6f);
    if (it > 0) {
      const bool pert_ok = (float)st.pert_pose_norm < ps.sp.d_pert;
      const float ds = score - ps.prev_score;
      const bool score_ok = (ds > 0.0f) && ds / (ps.prev_score + 1e-6f) < ps.sp.d_score;
      if (pert_ok && score_ok) stop = true;
    }
    ps.prev_score = score;
    if (stop) ps.status = 3;
  }
  if (stop && ps.status == 0) ps.status = 1;
  // next iteration's FP32 cache (updateCache, icp_impl_.h:16-21), formed in
  // FP64 and rounded once.
  IsoD Xc = iso_from_34<double>(ps.X);
  IsoD Xf = iso_from_34<double>(ps.Xf);
  ps.cache.Xm = iso_cast<float>(Xc);
  ps.cache.Xf = iso_cast<float>(Xf);
  ps.cache.Xmf = iso_cast<float>(compose(inverse(Xf), Xc));
  for (int i = 0; i < 6; ++i) ps.cache.v[i] = (float)ps.v[i];
  ps.cache.c = ps.sp.recenter ? apply_iso(ps.cache.Xm, ps.moving_root_mean) : v3(0.0f, 0.0f, 0.0f);
  ps.tiles_done = 0;
}

// ---- the tile kernel --------------------------------------------------------

template <bool CT, int LDS_NODES>
__global__ __launch_bounds__(kThreads) void tiles_kernel(PairState* __restrict__ pairs,
                                                         const int* __restrict__ tile_pair,
                                                         int n_tiles_total,
                                                         float* __restrict__ partials,
                                                         int* __restrict__ corr_out,
                                                         int* __restrict__ depth_out,
                                                         int xcd_remap) {
  constexpr int N = Slots<CT>::N;
  constexpr int P = Slots<CT>::P;
  constexpr int PL = P / 64;  // slots held per lane after the wave reduce
  __shared__ float4 s_nodes[LDS_NODES > 0 ? 2 * LDS_NODES : 2];
  __shared__ float s_wave[kWaves][P];
  __shared__ double s_red[N];
  __shared__ int s_last;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  int cached_pair = -1;
  int n_lds = 0;

  // Tile schedule.  Workgroups are dispatched round-robin over the 8 XCDs of
  // the chip, each with its own L2; with xcd_remap the XCD of workgroup b
  // (b % 8) walks a contiguous eighth of the tile list, so consecutive tiles of
  // one alignment -- which read the same fixed tree -- share an L2.
  const int G = gridDim.x;
  const bool remap = xcd_remap && (G % 8) == 0 && G >= 8;
  const int per_xcd = remap ? (n_tiles_total + 7) / 8 : 0;
  const int xcd = blockIdx.x & 7;
  const int j0 = remap ? (int)(blockIdx.x >> 3) : (int)blockIdx.x;
  const int stride = remap ? G / 8 : G;
  for (int k = j0;; k += stride) {
    int t;
    if (remap) {
      if (k >= per_xcd) break;
      t = xcd * per_xcd + k;
      if (t >= n_tiles_total) break;
    } else {
      if (k >= n_tiles_total) break;
      t = k;
    }
    const int p = tile_pair ? tile_pair[t] : 0;
    PairState& ps = pairs[p];
    if (ps.status != 0) continue;
    const DevTree fixed = ps.fixed;
    if (LDS_NODES > 0 && p != cached_pair) {
      __syncthreads();
      n_lds = fixed.n_nodes < LDS_NODES ? fixed.n_nodes : LDS_NODES;
      for (int i = tid; i < 2 * n_lds; i += kThreads) s_nodes[i] = fixed.nodes[i];
      cached_pair = p;
      __syncthreads();
    }
    const int lt = t - ps.tile_first;
    const int q = lt * kTile + tid;

    float v[P];
#pragma unroll
    for (int i = N; i < P; ++i) v[i] = 0.0f;
    int leaf = -1, depth = 0;
    if (q < ps.moving.n_leaves && fixed.n_nodes > 0) {
      const float4 la = ps.moving.leaves[2 * q];
      const float4 lb = ps.moving.leaves[2 * q + 1];
      LeafIn L;
      L.mean = v3(la.x, la.y, la.z);
      L.dir = v3(la.w, lb.x, lb.y);
      L.dts = lb.z;
      const TermCache C = ps.cache;
      const TermParams TP = ps.tp;
      if (LDS_NODES > 0) {
        LdsFetch f{fixed.nodes, s_nodes, n_lds};
        leaf = CT ? ct_term(f, C, TP, L, v, &depth) : icp_term(f, C, TP, L, v, &depth);
      } else {
        GlobalFetch f{fixed.nodes};
        leaf = CT ? ct_term(f, C, TP, L, v, &depth) : icp_term(f, C, TP, L, v, &depth);
      }
      if (corr_out) corr_out[q] = leaf;
      if (depth_out) depth_out[q] = depth;
    } else {
#pragma unroll
      for (int i = 0; i < N; ++i) v[i] = 0.0f;
    }

    wave_reduce_scatter<P>(v, lane);
#pragma unroll
    for (int j = 0; j < PL; ++j) s_wave[wave][j + PL * lane] = v[j];
    __syncthreads();
    if (tid < N) {
      const float s = (s_wave[0][tid] + s_wave[1][tid]) + (s_wave[2][tid] + s_wave[3][tid]);
      partials[(size_t)t * N + tid] = s;
    }
    __threadfence();
    __syncthreads();
    if (tid == 0) {
      const unsigned prev = atomicAdd(&ps.tiles_done, 1u);
      s_last = (prev == (unsigned)(ps.n_tiles - 1)) ? 1 : 0;
    }
    __syncthreads();
    if (s_last) {
      __threadfence();
      finish_pair<CT>(ps, partials, s_red);
      __syncthreads();
    }
  }
}

// ---- launch wrappers --------------------------------------------------------

template <bool CT, int LDS>
static hipError_t launch_t(const LaunchArgs& a, hipStream_t s) {
  hipLaunchKernelGGL((tiles_kernel<CT, LDS>), dim3(a.grid), dim3(kThreads), 0, s, a.pairs, a.tile_pair,
                     a.n_tiles_total, a.partials, a.corr_out, a.depth_out, a.xcd_remap);
  return hipGetLastError();
}

hipError_t launch_tiles(const LaunchArgs& a, hipStream_t s) {
  if (a.ct) {
    switch (a.lds_nodes) {
      case 0: return launch_t<true, 0>(a, s);
      case 256: return launch_t<true, 256>(a, s);
      case 512: return launch_t<true, 512>(a, s);
      default: return launch_t<true, 1024>(a, s);
    }
  }
  switch (a.lds_nodes) {
    case 0: return launch_t<false, 0>(a, s);
    case 256: return launch_t<false, 256>(a, s);
    case 512: return launch_t<false, 512>(a, s);
    case 1024: return launch_t<false, 1024>(a, s);
    default: return launch_t<false, 2048>(a, s);
  }
}

int slots(int ct) { return ct ? kCtSlots : kIcpSlots; }

}  // namespace kdh
