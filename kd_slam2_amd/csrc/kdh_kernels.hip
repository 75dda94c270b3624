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
//   2. evaluates the point-to-plane term (kdh_term.h) into a compact 16-value
//      record per lane,
//   3. expands the 34 (ICP) or 97 (CT) accumulator slots from the record in
//      chunks of 64 and reduces each chunk across the 64 lanes of the wave
//      with a butterfly reduce-scatter (63 shuffles per chunk instead of
//      64 x 6), then across the 4 waves through LDS,
//   4. writes one FP32 partial row per tile to a fixed slot.
// A second kernel, `finish_kernel`, runs one workgroup per alignment: it sums
// the alignment's partial rows in FP64 in tile order -- the result does not
// depend on which workgroup ran which tile, nor on what else shares the
// launch -- and then either publishes H/b (build mode) or runs the FP64
// damped LDLT solve, the SE(3) update and the next iteration's FP32 cache
// (solve mode).  Keeping the FP64 solve out of the tile kernel keeps the
// tile kernel's register budget small enough for several waves per SIMD.
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

// Butterfly reduce-scatter over one wave: on entry every lane holds 64
// values; after the step with lane offset OFF the lane keeps OFF values (the
// upper or lower half, by its OFF bit) and adds its partner's other half.
// On exit lane l holds the wave sum of slot l in v[0].  The pairing is fixed,
// so the summation order -- and the result -- is deterministic.
template <int OFF>
__device__ __forceinline__ void reduce_scatter_step(float* v, int lane) {
  const bool upper = (lane & OFF) != 0;
#pragma unroll
  for (int i = 0; i < OFF; ++i) {
    const float send = upper ? v[i] : v[i + OFF];
    const float keep = upper ? v[i + OFF] : v[i];
    v[i] = keep + __shfl_xor(send, OFF, 64);
  }
}

__device__ __forceinline__ void wave_reduce_scatter64(float* v, int lane) {
  reduce_scatter_step<32>(v, lane);
  reduce_scatter_step<16>(v, lane);
  reduce_scatter_step<8>(v, lane);
  reduce_scatter_step<4>(v, lane);
  reduce_scatter_step<2>(v, lane);
  reduce_scatter_step<1>(v, lane);
}

struct GlobalFetch {
  const float4* __restrict__ nodes;
  __device__ __forceinline__ void operator()(int idx, float4& a, float4& b) const {
    a = nodes[2 * idx];
    b = nodes[2 * idx + 1];
  }
};

// The LDS copy is addressed through an explicit LDS (address space 3)
// pointer: with a generic one the backend must select between the two
// apertures per load.
#if defined(__HIP_DEVICE_COMPILE__)
typedef __attribute__((address_space(3))) const float4 lds_f4;
#else
typedef const float4 lds_f4;
#endif
struct LdsFetch {
  const float4* __restrict__ nodes;
  lds_f4* lds;
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

// ---- per-pair final stage (finish_kernel) ----------------------------------

// Broadcast of a double from lane `src` (wave-uniform) through v_readlane:
// a scalar read of one lane's register, much shorter latency than an LDS
// permute, which matters on the serial chain of the factorisation.
__device__ __forceinline__ double bcast(double x, int src) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(x), src);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(x), src);
  return __hiloint2double(hi, lo);
}

// Damped solve (A + diag(damp)) x = -b on one wave: lane i (< D) holds row i
// of the symmetric matrix in registers; right-looking LDL^T without pivoting
// (the damped normal matrix is positive definite), then forward and backward
// substitution, all with lane broadcasts.  Lanes >= D carry zeros.
template <int D>
__device__ void wave_ldlt_solve(const double* sH, const double* sb, double damp_pose, double damp_vel, int lane,
                                double* sx) {
  double a[D];
  const bool row = lane < D;
#pragma unroll
  for (int j = 0; j < D; ++j) a[j] = row ? sH[lane * D + j] : 0.0;
  if (row) {
#pragma unroll
    for (int j = 0; j < D; ++j)
      if (j == lane) a[j] += (j < 6) ? damp_pose : damp_vel;
  }
  double r = row ? -sb[lane] : 0.0;
  double dd = 0.0;  // this lane's pivot D_i
#pragma unroll
  for (int k = 0; k < D; ++k) {
    double akk = 0.0;
#pragma unroll
    for (int j = 0; j < D; ++j)
      if (j == k) akk = a[j];
    const double dk = bcast(akk, k);
    if (lane == k) dd = dk;
    const double inv = (dk != 0.0) ? 1.0 / dk : 0.0;
    const double lik = (lane > k && row) ? a[k] * inv : 0.0;
    // A_ij -= L_ik * A_kj for j > k (row k broadcast from lane k)
#pragma unroll
    for (int j = k + 1; j < D; ++j) {
      const double akj = bcast(a[j], k);
      a[j] -= lik * akj;
    }
    if (lane > k && row) a[k] = lik;  // store L_ik in the strictly lower part
  }
  // forward: L y = r
#pragma unroll
  for (int k = 0; k < D; ++k) {
    const double yk = bcast(r, k);
    if (lane > k && row) r -= a[k] * yk;
  }
  r = (row && dd != 0.0) ? r / dd : 0.0;
  // backward: L^T x = z; lane i needs L_ki = a[i] of lane k
#pragma unroll
  for (int k = D - 1; k >= 0; --k) {
    const double xk = bcast(r, k);
#pragma unroll
    for (int i = 0; i < k; ++i) {
      const double lki = bcast(a[i], k);
      if (lane == i) r -= lki * xk;
    }
  }
  if (row) sx[lane] = r;
}

// LDS needed by finish_pair (bytes); the caller passes an area at least this
// large, 16-byte aligned.
template <bool CT>
struct FinishLds {
  static constexpr int N = Slots<CT>::N;
  static constexpr int D = CT ? 12 : 6;
  static constexpr int CH = 32;  // tiles staged per round
  static constexpr int kRed = 0;                          // double[N]
  static constexpr int kH = kRed + 8 * ((N + 1) & ~1);    // double[D*D]
  static constexpr int kG = kH + 8 * D * D;               // double[D*D]
  static constexpr int kB = kG + 8 * D * D;               // double[D]
  static constexpr int kX = kB + 8 * D;                   // double[D]
  static constexpr int kGo = kX + 8 * D;                  // int (padded)
  static constexpr int kStage = kGo + 16;                 // float[CH*N]
  static constexpr int kBytes = kStage + 4 * CH * N;
};

template <bool CT>
__device__ void finish_pair(PairState& ps, const float* __restrict__ partials, char* lds, int tile_first,
                            int* status_out) {
  using FL = FinishLds<CT>;
  constexpr int N = Slots<CT>::N;
  constexpr int D = CT ? 12 : 6;
  constexpr int CS = CT ? 90 : 27;  // chi2 slot
  constexpr int CH = FL::CH;
  float* s_stage = (float*)(lds + FL::kStage);
  double* s_red = (double*)(lds + FL::kRed);
  double* sH = (double*)(lds + FL::kH);
  double* sG = (double*)(lds + FL::kG);
  double* sb = (double*)(lds + FL::kB);
  double* sx = (double*)(lds + FL::kX);
  int& s_go = *(int*)(lds + FL::kGo);
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  // Read the whole pair state the stage needs up front (uniform addresses:
  // scalar loads issued together), so its latency overlaps the partial-row
  // loads instead of adding round trips at the end.
  const int mode = ps.mode;
  const int it = ps.iter;
  const int stats_mode = ps.tp.stats_mode;
  const SolveParams sp = ps.sp;
  const bool has_prior = ps.has_prior != 0.0;
  const V3f root_mean = ps.moving_root_mean;
  const float prev_score = ps.prev_score;
  double Xs[12], Xfs[12], vs[6];
#pragma unroll
  for (int i = 0; i < 12; ++i) {
    Xs[i] = ps.X[i];
    Xfs[i] = ps.Xf[i];
  }
#pragma unroll
  for (int i = 0; i < 6; ++i) vs[i] = ps.v[i];

  // 1. FP64 sum of the pair's partial rows in tile order.  Rows are staged
  //    through LDS CH at a time with coalesced loads by all threads; thread s
  //    (< N) adds them in order, so the result is deterministic.
  const float* base = partials + (size_t)tile_first * N;
  double acc = 0.0;
  for (int t0 = 0; t0 < ps.n_tiles; t0 += CH) {
    const int nt = min(CH, ps.n_tiles - t0);
    __syncthreads();
    for (int i = tid; i < nt * N; i += kThreads) s_stage[i] = base[(size_t)t0 * N + i];
    __syncthreads();
    if (tid < N)
      for (int t = 0; t < nt; ++t) acc += (double)s_stage[t * N + tid];
  }
  if (tid < N) s_red[tid] = acc;
  __syncthreads();

  // 2. Assemble the re-centred H_c, b_c, then map back: G = H_c M, H = M^T G,
  //    b = M^T b_c with M = [[I, -[c]x], [0, I]] (see uncenter_d).
  const V3d cc = v3(ps.c64[0], ps.c64[1], ps.c64[2]);
  const M3d Cx = skew(cc);
  for (int e = tid; e < D * D; e += kThreads) {
    const int r = e / D, c = e % D;
    double h;
    if (r < 6 && c < 6) {
      const int rr = r < c ? r : c, c2 = r < c ? c : r;
      const int k = rr * 6 - (rr * (rr - 1)) / 2 + (c2 - rr);
      h = s_red[k];
    } else if (r >= 6 && c >= 6) {
      const int rr = (r < c ? r : c) - 6, c2 = (r < c ? c : r) - 6;
      const int k = rr * 6 - (rr * (rr - 1)) / 2 + (c2 - rr);
      h = s_red[27 + k];
    } else if (r < 6) {
      h = s_red[54 + r * 6 + (c - 6)];
    } else {
      h = s_red[54 + c * 6 + (r - 6)];
    }
    sH[e] = h;
  }
  if (tid < D) sb[tid] = s_red[tid < 6 ? 21 + tid : 48 + (tid - 6)];
  __syncthreads();
  for (int e = tid; e < D * D; e += kThreads) {
    const int r = e / D, c = e % D;
    double g = sH[e];
    if (c >= 3 && c < 6)
      g -= (sH[r * D + 0] * Cx(0, c - 3) + sH[r * D + 1] * Cx(1, c - 3)) + sH[r * D + 2] * Cx(2, c - 3);
    sG[e] = g;
  }
  __syncthreads();
  for (int e = tid; e < D * D; e += kThreads) {
    const int r = e / D, c = e % D;
    double h = sG[e];
    if (r >= 3 && r < 6)
      h -= (Cx(0, r - 3) * sG[0 * D + c] + Cx(1, r - 3) * sG[1 * D + c]) + Cx(2, r - 3) * sG[2 * D + c];
    sH[e] = h;
  }
  double bj = 0.0;
  if (tid >= 3 && tid < 6)
    bj = sb[tid] - ((Cx(0, tid - 3) * sb[0] + Cx(1, tid - 3) * sb[1]) + Cx(2, tid - 3) * sb[2]);
  __syncthreads();
  if (tid >= 3 && tid < 6) sb[tid] = bj;
  __syncthreads();

  DevStats st;
  st.num_inliers = (int)s_red[CS + 2 + kInlier];
  st.num_outliers = (int)s_red[CS + 2 + kOutlier];
  st.num_bad = (int)s_red[CS + 2 + kBad];
  st.num_err_normal = (int)s_red[CS + 2 + kErrNormal];
  st.num_err_fn = (int)s_red[CS + 2 + kErrFn];
  st.pad = 0;
  st.chi2 = s_red[CS];
  st.chi2_kernelized = s_red[CS + 1];
  st.pert_pose_norm = 0.0;
  st.pert_vel_norm = 0.0;

  if (mode == 1) {  // build only: publish H, b, stats
    for (int e = tid; e < D * D; e += kThreads) ps.H[e] = sH[e];
    if (tid < D) ps.b[tid] = sb[tid];
    if (tid == 0) ps.stats_log[0] = st;
    return;
  }

  // 3. pose prior (tree_aligner_base_.h:139-152), serial on thread 0
  if (has_prior && tid == 0) {
    IsoD X = iso_from_34<double>(Xs);
    IsoD ZXi = iso_from_34<double>(ps.prior_ZXi);
    double e[6], J[36], OJ[36], Oe[6];
    logmap_d(compose(ZXi, X), e);
    dPoseToPoseError_d(ZXi, X, J);
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
      double sbv = 0.0;
      for (int k = 0; k < 6; ++k) sbv += J[k * 6 + r] * Oe[k];
      sb[r] += sbv;
      for (int c = 0; c < 6; ++c) {
        double s = 0.0;
        for (int k = 0; k < 6; ++k) s += J[k * 6 + r] * OJ[k * 6 + c];
        sH[r * D + c] += s;
      }
    }
  }
  __syncthreads();
  for (int e = tid; e < D * D; e += kThreads) ps.H[e] = sH[e];
  if (tid < D) ps.b[tid] = sb[tid];

  // 4. min_good gate (icp_impl_.h:76-80)
  if (tid == 0) s_go = (st.num_inliers + st.num_outliers >= sp.min_good_measurements) ? 1 : 0;
  __syncthreads();
  if (!s_go) {
    if (tid == 0) {
      ps.stats_log[it] = st;
      ps.iter = it + 1;
      ps.status = 2;  // failed: compute() returns -it
      if (status_out) status_out[0] = 2;
    }
    return;
  }

  // 5. damped solve on wave 0
  if (tid < 64) wave_ldlt_solve<D>(sH, sb, sp.damping, sp.velocity_damping, lane, sx);
  __syncthreads();
  if (tid != 0) return;

  // 6. update, stats, termination, next cache (thread 0)
  double dx[D];
  for (int i = 0; i < D; ++i) dx[i] = sx[i];
  double pn = 0.0, vn = 0.0;
  for (int i = 0; i < 6; ++i) pn += dx[i] * dx[i];
  if (CT)
    for (int i = 6; i < D; ++i) vn += dx[i] * dx[i];
  st.pert_pose_norm = sqrt(pn);
  st.pert_vel_norm = sqrt(vn);
  ps.stats_log[it] = st;
  ps.iter = it + 1;
  IsoD Xc = iso_from_34<double>(Xs);
  if (!stats_mode) {
    Xc = compose(expmap_d(dx), Xc);
    iso_to_34(Xc, ps.X);
    if (CT)
      for (int i = 0; i < 6; ++i) {
        vs[i] += dx[6 + i];
        ps.v[i] = vs[i];
      }
  }
  // termination: stats mode runs one round; the score criterion
  // (icp_stats_.h:85-102) may stop early; otherwise max_iterations.
  bool stop = stats_mode != 0 || it + 1 >= sp.max_iterations;
  int status = 0;
  if (!stop && sp.use_term_crit) {
    const float ni = (float)st.num_inliers;
    const int nl = st.num_inliers + st.num_outliers + st.num_bad + st.num_err_normal + st.num_err_fn;
    const float ratio = nl ? ni / (float)nl : 0.0f;
    const float score = ratio / ((float)st.chi2_kernelized / ni + 1e-6f);
    if (it > 0) {
      const bool pert_ok = (float)st.pert_pose_norm < sp.d_pert;
      const float ds = score - prev_score;
      const bool score_ok = (ds > 0.0f) && ds / (prev_score + 1e-6f) < sp.d_score;
      if (pert_ok && score_ok) stop = true;
    }
    ps.prev_score = score;
    if (stop) status = 3;
  }
  if (stop && status == 0) status = 1;
  ps.status = status;
  if (status_out) status_out[0] = status;
  // next iteration's FP32 cache (updateCache, icp_impl_.h:16-21), formed in
  // FP64 and rounded once.
  const IsoD Xf = iso_from_34<double>(Xfs);
  TermCache nc;
  nc.Xm = iso_cast<float>(Xc);
  nc.Xf = iso_cast<float>(Xf);
  nc.Xmf = iso_cast<float>(compose(inverse(Xf), Xc));
  for (int i = 0; i < 6; ++i) nc.v[i] = (float)vs[i];
  const Anchor an = make_anchor(Xc, Xf, root_mean, sp.recenter);
  nc.mroot = an.mroot;
  nc.cf = an.cf;
  ps.cache = nc;
  ps.c64[0] = an.c.x;
  ps.c64[1] = an.c.y;
  ps.c64[2] = an.c.z;
}

// ---- the tile kernel --------------------------------------------------------

template <bool CT, int LDS_NODES>
__global__ __launch_bounds__(kThreads) void tiles_kernel(const PairState* __restrict__ pairs,
                                                         const int* __restrict__ tile_pair,
                                                         int n_tiles_total,
                                                         float* __restrict__ partials,
                                                         int* __restrict__ corr_out,
                                                         int* __restrict__ depth_out,
                                                         int xcd_remap,
                                                         const int* __restrict__ pair_tile_first) {
  constexpr int N = Slots<CT>::N;
  constexpr int NCH = (N + 63) / 64;  // 64-slot chunks
  __shared__ float4 s_nodes[LDS_NODES > 0 ? 2 * LDS_NODES : 2];
  __shared__ float s_wave[kWaves][NCH * 64];

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
    const PairState& ps = pairs[p];
    if (ps.status != 0) continue;
    const DevTree fixed = ps.fixed;
    if (LDS_NODES > 0 && p != cached_pair) {
      __syncthreads();
      n_lds = fixed.n_nodes < LDS_NODES ? fixed.n_nodes : LDS_NODES;
      for (int i = tid; i < 2 * n_lds; i += kThreads) s_nodes[i] = fixed.nodes[i];
      cached_pair = p;
      __syncthreads();
    }
    const int lt = t - (pair_tile_first ? pair_tile_first[p] : ps.tile_first);
    const int q = lt * kTile + tid;

    TermRec R;
    rec_reject(R, -1);  // padding lanes contribute nothing
    if (q < ps.moving.n_leaves && fixed.n_nodes > 0) {
      const float4 la = ps.moving.leaves[2 * q];
      const float4 lb = ps.moving.leaves[2 * q + 1];
      LeafIn L;
      L.mean = v3(la.x, la.y, la.z);
      L.dir = v3(la.w, lb.x, lb.y);
      L.dts = lb.z;
      const TermCache C = ps.cache;
      const TermParams TP = ps.tp;
      int leaf, depth = 0;
      if (LDS_NODES > 0) {
        LdsFetch f{fixed.nodes, (lds_f4*)s_nodes, n_lds};
        leaf = CT ? ct_term(f, C, TP, L, R, &depth) : icp_term(f, C, TP, L, R, &depth);
      } else {
        GlobalFetch f{fixed.nodes};
        leaf = CT ? ct_term(f, C, TP, L, R, &depth) : icp_term(f, C, TP, L, R, &depth);
      }
      if (corr_out) corr_out[q] = leaf;
      if (depth_out) depth_out[q] = depth;
    }

    // Expand slots chunk by chunk and reduce each over the wave.
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch) {
      float v[64];
#pragma unroll
      for (int i = 0; i < 64; ++i) v[i] = rec_slot<CT>(R, ch * 64 + i);
      wave_reduce_scatter64(v, lane);
      s_wave[wave][ch * 64 + lane] = v[0];
    }
    __syncthreads();
    if (tid < N) {
      const float s = (s_wave[0][tid] + s_wave[1][tid]) + (s_wave[2][tid] + s_wave[3][tid]);
      partials[(size_t)t * N + tid] = s;
    }
    __syncthreads();
  }
}

// One workgroup per alignment: FP64 reduction of its partial rows, then
// build-mode publish or solve-mode Gauss-Newton step.
template <bool CT>
__global__ __launch_bounds__(kThreads) void finish_kernel(PairState* __restrict__ pairs, int pair_first,
                                                          int n_pairs, const float* __restrict__ partials,
                                                          const int* __restrict__ pair_tile_first,
                                                          int* __restrict__ status_out) {
  __shared__ __align__(16) char s_area[FinishLds<CT>::kBytes];
  const int p = pair_first + blockIdx.x;
  if (blockIdx.x >= n_pairs) return;
  PairState& ps = pairs[p];
  if (ps.status != 0 || ps.ct != (CT ? 1 : 0) || ps.n_tiles == 0) return;
  finish_pair<CT>(ps, partials, s_area, pair_tile_first ? pair_tile_first[p] : ps.tile_first,
                  status_out ? status_out + p : nullptr);
}

// ---- launch wrappers --------------------------------------------------------

template <bool CT, int LDS>
static hipError_t launch_t(const LaunchArgs& a, hipStream_t s) {
  if (a.phases & 1) {
    hipLaunchKernelGGL((tiles_kernel<CT, LDS>), dim3(a.grid), dim3(kThreads), 0, s, a.pairs, a.tile_pair,
                       a.n_tiles_total, a.partials, a.corr_out, a.depth_out, a.xcd_remap, a.pair_tile_first);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if ((a.phases & 2) && a.n_pairs > 0) {
    hipLaunchKernelGGL((finish_kernel<CT>), dim3(a.n_pairs), dim3(kThreads), 0, s, a.pairs, a.pair_first,
                       a.n_pairs, a.partials, a.pair_tile_first, a.status_out);
    return hipGetLastError();
  }
  return hipSuccess;
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
