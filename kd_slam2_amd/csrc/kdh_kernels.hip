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
//   2. evaluates the point-to-plane term (kdh_term.h) into a compact record
//      per lane (Jacobians, residual, square root of the robust weight,
//      outcome),
//   3. forms the tile's normal equations on the matrix cores: every
//      accumulator slot is an entry of C = sum_leaves w w^T over the leaf
//      vectors w = [sqrt(g) J, sqrt(g) Jv, sqrt(g) e, e], one 16x16 C per
//      wave from v_mfma_f32_16x16x4f32 with the 64 leaves as the K dimension
//      (operands transposed through LDS; the rigid form's 8-component
//      vectors fill both halves of the 16 rows with two leaf sets, 8 steps
//      per wave instead of 16), the outcome counts from ballots, then the
//      4 waves' C tiles summed through LDS in a fixed order,
//   4. writes one FP32 partial row (34 ICP / 97 CT slots) per tile to a
//      fixed slot.
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

template <bool CT> struct Slots;
template <> struct Slots<false> { static constexpr int N = kIcpSlots; static constexpr int P = 64; };
template <> struct Slots<true> { static constexpr int N = kCtSlots; static constexpr int P = 128; };
// The dual form appends Diag(vel_B) 27, Cross(pose, vel_B) 36 and
// Cross(vel_A, vel_B) 36 (icp_ws_item_.h) to the 97 CT slots.
constexpr int kDualSlots = kCtSlots + 99;
template <bool CT, bool DUAL> constexpr int n_slots() { return DUAL ? kDualSlots : Slots<CT>::N; }

typedef float floatx4 __attribute__((ext_vector_type(4)));

// Orders this wave's LDS writes before its later LDS reads (one wave: LDS
// executes its instructions in order; the fences keep the compiler from
// moving them).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Dual extras (slot s >= 97): cell of C_zz = sum z z^T (tile 2) or of
// C_uz = sum w z^T (tile 1), z = sqrt(g) J_vel_B; returned as tile * 256 + cell.
__device__ __forceinline__ int dual_slot_cell(int s) {
  const int t = s - kCtSlots;
  if (t < 21) return 2 * 256 + tri_c(t) * 16 + tri_r(t);  // Diag(vel_B) J_b J_b^T g
  if (t < 27) return 256 + (t - 21) * 16 + 12;             // J_b e g: row sqrt(g) e of w
  if (t < 63) return 256 + ((t - 27) % 6) * 16 + (t - 27) / 6;        // J_pose J_b^T g
  return 256 + ((t - 63) % 6) * 16 + 6 + (t - 63) / 6;                 // J_vel J_b^T g
}

// Cell (j * 16 + i) of C = sum w w^T holding accumulator slot s < CS + 2, in
// the packing of ICPWsItemDiag_::pack (upper-triangular J J^T g, then J e g;
// icp_ws_item_.h:26-34) and ICPWsItemCross_::pack (J_pose J_vel^T g,
// icp_ws_item_.h:72-75), then chi2 and chi2 kernelised.
template <bool CT>
__device__ __forceinline__ int slot_cell(int s) {
  constexpr int E = CT ? 12 : 6;
  constexpr int CS = CT ? 90 : 27;
  int i, j;
  if (s < 21) {
    i = tri_r(s);
    j = tri_c(s);
  } else if (s < 27) {
    i = s - 21;
    j = E;
  } else if (CT && s < 48) {
    i = 6 + tri_r(s - 27);
    j = 6 + tri_c(s - 27);
  } else if (CT && s < 54) {
    i = 6 + (s - 48);
    j = E;
  } else if (CT && s < 90) {
    i = (s - 54) / 6;
    j = 6 + (s - 54) % 6;
  } else if (s == CS) {
    i = j = E + 1;  // chi2 = sum e^2
  } else {
    i = j = E;      // chi2 kernelised = sum g e^2
  }
  return j * 16 + i;
}

// Tree records are read through explicit global (address space 1) pointers
// and the LDS copy through explicit LDS (address space 3) pointers.  Pointers
// loaded from the pair state are generic: through them every load would be a
// flat load, which the backend must route between the apertures and which
// counts against the LDS wait counter too, so each wait for a tree record
// would also wait for every LDS access in flight.
// Staged records travel as clang's own 4-float vector: arrays of the HIP
// float4 class that are written under a branch stay in scratch memory.
typedef float v4f __attribute__((ext_vector_type(4)));
#if defined(__HIP_DEVICE_COMPILE__)
typedef __attribute__((address_space(3))) const float4 lds_f4;
typedef __attribute__((address_space(1))) const float4 glb_f4;
typedef __attribute__((address_space(1))) const float glb_f;
typedef __attribute__((address_space(1))) const v4f glb_v4;
typedef __attribute__((address_space(3))) v4f lds_v4;
#else
typedef const float4 lds_f4;
typedef const float4 glb_f4;
typedef const float glb_f;
typedef const v4f glb_v4;
typedef v4f lds_v4;
#endif
__device__ __forceinline__ glb_f4* global_f4(const float4* p) { return (glb_f4*)p; }
// A valid address for loads whose value is never used (a tree without nodes).
__device__ const float4 kNoRecord[1] = {{0.0f, 0.0f, 0.0f, 0.0f}};
__device__ __forceinline__ glb_f* global_f(const float* p) { return (glb_f*)p; }

// A tile record in one 32-byte load (scalar for a wave-uniform address):
// a plain struct copy may be split into field loads, some of which the
// compiler then sinks below later branches.
__device__ __forceinline__ TileDesc load_tile_desc(const TileDesc* p) {
  typedef unsigned int u8v __attribute__((ext_vector_type(8)));
  const u8v v = *reinterpret_cast<const u8v*>(p);
  TileDesc d;
  d.nodes = reinterpret_cast<const float4*>(((unsigned long long)v[1] << 32) | v[0]);
  d.leaves = reinterpret_cast<const float4*>(((unsigned long long)v[3] << 32) | v[2]);
  d.n_nodes = (int)v[4];
  d.n_leaves = (int)v[5];
  d.pair = (int)v[6];
  d.leaf0 = (int)v[7];
  return d;
}

struct GlobalFetch {
  glb_f4* __restrict__ nodes;
  __device__ __forceinline__ void operator()(int idx, float4& a, float4& b) const {
    a = nodes[2 * idx];
    b = nodes[2 * idx + 1];
  }
  __device__ __forceinline__ int descend(V3f q, int* depth_out, float4& a, float4& b) const {
    return kdh::descend(*this, q, depth_out, a, b);
  }
};

// One greedy step at the node whose record is (a, b) (the step of `descend`,
// kdh_term.h): false at a leaf (idx stays the leaf); otherwise idx becomes the
// child on q's side, -1 when that child is missing.
__device__ __forceinline__ bool descent_step(V3f q, const float4& a, const float4& b, int& idx) {
  const int l = __float_as_int(b.z);
  const int r = __float_as_int(b.w);
  if ((l & r) < 0) return false;  // both links negative: a leaf
  const float dx = q.x - a.x;
  const float dy = q.y - a.y;
  const float dz = q.z - a.z;
  const float s = (a.w * dx + b.x * dy) + b.y * dz;
  idx = s > 0.0f ? l : r;
  return true;
}

// Records 0 .. n_lds-1 (the top of the tree in BFS order) come from the LDS
// copy, the rest from global memory.  DEPTH: the descent counts its steps
// into *depth_out (the launch asked for depths).
template <bool DEPTH>
struct LdsFetch {
  glb_f4* __restrict__ nodes;
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
  // The greedy descent in two loops: the staged levels from LDS until the
  // path leaves them, then the rest from global memory.  A BFS prefix is
  // closed under parents, so once a path leaves it, it never comes back, and
  // each loop tests only what its records need: the decisions and their order
  // are those of `descend`.
  __device__ __forceinline__ int descend(V3f q, int* depth_out, float4& a, float4& b) const {
    int idx = 0;
    int depth = 0;
    bool deep = true;  // the descent goes on at idx in global memory
    if (n_lds > 0) {
      a = lds[0];
      b = lds[1];
      for (;;) {
        if (!descent_step(q, a, b, idx)) {
          deep = false;
          break;
        }
        if (DEPTH) ++depth;
        if (idx < 0) {
          deep = false;
          break;
        }
        if (idx >= n_lds) break;
        a = lds[2 * idx];
        b = lds[2 * idx + 1];
      }
    }
    if (deep) {
      a = nodes[2 * idx];
      b = nodes[2 * idx + 1];
      for (;;) {
        if (!descent_step(q, a, b, idx)) break;
        if (DEPTH) ++depth;
        if (idx < 0) break;
        a = nodes[2 * idx];
        b = nodes[2 * idx + 1];
      }
    }
    if (DEPTH) *depth_out = depth;
    return idx;
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

// 1/d to full double precision: the hardware reciprocal estimate refined by
// two Newton steps -- half the dependent operations of an IEEE division on
// the factorisation's serial pivot chain (not correctly rounded).
__device__ __forceinline__ double rcp_nr(double d) {
  double y = __builtin_amdgcn_rcp(d);
  double e = fma(-d, y, 1.0);
  y = fma(y, e, y);
  e = fma(-d, y, 1.0);
  return fma(y, e, y);
}

// Damped solve (A + diag(damp)) x = -b on one wave: lane i (< D) holds row i
// of the symmetric matrix in a[] and b_i in bi; right-looking elimination
// without pivoting (the damped normal matrix is positive definite) with the
// forward substitution carried along, then backward substitution, with lane
// broadcasts.  Lanes >= D carry zeros.  After the elimination lane i holds
// row i of U = D L^T in a[j], j > i; scaled by its own pivot reciprocal that
// is row i of L^T, so the backward steps need no transpose.  The caller's
// workgroup is this one wave.  The result is the same on the wave alone or
// in any batch; it differs from the host solve of kdh_solve.h in the last
// bits (fused multiply-adds, order of the back-substitution products).
template <int D>
__device__ void wave_ldlt_solve(double (&a)[D], double bi, double damp_pose, double damp_vel, int lane,
                                double* sx) {
  // FP64 only, and no decision rests on its last bits: multiply-adds may fuse
  // here (one dependent operation per update instead of two), unlike the
  // FP32 code whose rounding the decisions follow.
#pragma clang fp contract(fast)
  const bool row = lane < D;
  // the damping on the diagonal, by selects (a branch per column would split
  // the wave lane by lane); off the diagonal the added zero changes nothing
#pragma unroll
  for (int j = 0; j < D; ++j) a[j] = a[j] + ((row && j == lane) ? ((j < 6) ? damp_pose : damp_vel) : 0.0);
  double r = row ? -bi : 0.0;
  double my_inv = 0.0;  // this lane's pivot reciprocal 1 / D_i
#pragma unroll
  for (int k = 0; k < D; ++k) {
    double akk = 0.0;
#pragma unroll
    for (int j = 0; j < D; ++j)
      if (j == k) akk = a[j];
    const double dk = bcast(akk, k);
    const double rk = bcast(r, k);  // y_k: every update of row k is in
    const double inv = (dk != 0.0) ? rcp_nr(dk) : 0.0;
    if (lane == k) my_inv = inv;
    const double lik = (lane > k && row) ? a[k] * inv : 0.0;
    // A_ij -= L_ik * A_kj for j > k (row k broadcast from lane k), and the
    // same for the right-hand side
#pragma unroll
    for (int j = k + 1; j < D; ++j) {
      const double akj = bcast(a[j], k);
      a[j] -= lik * akj;
    }
    r -= lik * rk;
  }
  // z = D^-1 y, and row i of L^T = row i of U / D_i (lanes >= D: zeros)
  r *= my_inv;
#pragma unroll
  for (int j = 0; j < D; ++j) a[j] = j > lane ? a[j] * my_inv : 0.0;
  // backward: L^T x = z
#pragma unroll
  for (int k = D - 1; k >= 0; --k) {
    const double xk = bcast(r, k);
    if (lane < k) r -= a[k] * xk;
  }
  if (row) sx[lane] = r;
}

// LDS needed by finish_pair (bytes); the caller passes an area at least this
// large, 16-byte aligned.
template <bool CT>
struct FinishLds {
  static constexpr int N = kDualSlots;  // room for the largest (dual) row
  static constexpr int D = CT ? 12 : 6;
  static constexpr int kRed = 0;                          // double[N]
  static constexpr int kG = kRed + 8 * ((N + 1) & ~1);    // double[3*D]: rows 0..2 of G
  static constexpr int kP = kG + 8 * 3 * D;               // double[36 + 6]: prior H, b
  static constexpr int kX = kP + 8 * 42;                  // double[D]
  static constexpr int kL = kX + 8 * D;                   // double[40]: update scratch
  static constexpr int kBytes = kL + 8 * 40;
};

// Row r (< D) of H_c from the reduced slots, for a lane-dependent r and every
// column: the slot of (r, c) is a per-lane base plus a column offset, so the
// loop over c is one add and one LDS read per entry.  Entries (r, c) of a
// diagonal block with c < r come from the upper triangle (c, r), whose slot
// is a compile-time base plus r.
template <bool CT>
__device__ __forceinline__ void hc_row(const double* s_red, int r, double (&h)[CT ? 12 : 6]) {
  constexpr int D = CT ? 12 : 6;
  const bool pose = r < 6;
  const int rl = pose ? r : r - 6;                       // row within its diagonal block
  const int tri_r = rl * 6 - (rl * (rl - 1)) / 2 - rl;   // slot of (rl, c) is tri_r + c, c >= rl
  const int blk = pose ? 0 : 27;
#pragma unroll
  for (int c = 0; c < D; ++c) {
    const int cl = c < 6 ? c : c - 6;
    const int tri_c = cl * 6 - (cl * (cl - 1)) / 2 - cl;  // compile time: slot of (cl, rl) is tri_c + rl
    const bool same = pose == (c < 6);
    const int diag = blk + (cl >= rl ? tri_r + cl : tri_c + rl);
    const int cross = pose ? 54 + r * 6 + (c - 6) : 54 + c * 6 + rl;
    h[c] = s_red[same ? diag : cross];
  }
}

// Column j of [c]x (skew), chosen by selects: indexing a private matrix with
// a lane-dependent j would put it in scratch memory.
__device__ __forceinline__ V3d skew_col(const V3d& c, int j) {
  return j == 0 ? v3(0.0, c.z, -c.y) : j == 1 ? v3(-c.z, 0.0, c.x) : v3(c.y, -c.x, 0.0);
}

// The part of PairState the finish stage reads (inputs and state, up to the
// outputs), in 16-byte units.
constexpr int kStateVec4 = (int)((offsetof(PairState, stats_log) + 15) / 16);
static_assert(sizeof(PairState) % 16 == 0, "PairState rows must stay 16-byte aligned");

// One wave (64 lanes) per alignment.  Every step is lane-parallel or a
// one-wave chain; the workgroup barrier of a one-wave group costs nothing.
template <bool CT, bool DUAL>
__device__ void finish_pair(PairState& ps, const PairState& S, const float* __restrict__ partials, char* lds,
                            int tile_first, int* status_out) {
  using FL = FinishLds<CT>;
  constexpr int N = n_slots<CT, DUAL>();
  constexpr int D = CT ? 12 : 6;
  constexpr int CS = CT ? 90 : 27;  // chi2 slot
  double* s_red = (double*)(lds + FL::kRed);
  double* sG = (double*)(lds + FL::kG);
  double* sP = (double*)(lds + FL::kP);
  double* sx = (double*)(lds + FL::kX);
  double* sL = (double*)(lds + FL::kL);
  const int lane = threadIdx.x;
  const int tid = lane;
  // Inputs come from S, the LDS copy of the pair state taken at kernel start;
  // results go to ps in global memory.
  const int mode = S.mode;
  const int it = S.iter;
  const int nt = S.group_tiles;  // the tiles of every forest entry, contiguous
  const int n_entries = S.group_n;
  PairState* members = &ps;     // entries 1 .. n_entries - 1 follow the leader
  const int stats_mode = S.tp.stats_mode;
  const SolveParams sp = S.sp;
  const bool has_prior = S.has_prior != 0.0;
  const V3f root_mean = S.moving_root_mean;
  const float prev_score = S.prev_score;
  const V3d cc = v3(S.c64[0], S.c64[1], S.c64[2]);

  // 1. FP64 sum of the pair's partial rows in tile order: lane l owns slots
  //    l and l + 64 and adds their column straight from global memory.  All
  //    loads of a 32-row block are issued before the first add, so a block
  //    costs one memory round trip.  The order is fixed, so the result does
  //    not depend on the batch.
  // (The dual form's 196 slots take two such passes of 128.)
  for (int sb = 0; sb < N; sb += 128) {
    constexpr int TB = 32;
    const float* base = partials + (size_t)tile_first * N;
    const bool has0 = sb + lane < N, has1 = sb + lane + 64 < N;
    double acc0 = 0.0, acc1 = 0.0;
    for (int t0 = 0; t0 < nt; t0 += TB) {
      float v0[TB], v1[TB];
      // Unconditional loads from clamped, in-range addresses (nt >= 1): a
      // guarded load would become a branch with a wait after every load.
      const int s0 = has0 ? sb + lane : 0, s1 = has1 ? sb + lane + 64 : 0;
#pragma unroll
      for (int u = 0; u < TB; ++u) {
        const size_t row = (size_t)min(t0 + u, nt - 1) * N;
        v0[u] = base[row + s0];
        v1[u] = base[row + s1];
      }
#pragma unroll
      for (int u = 0; u < TB; ++u)
        if (t0 + u < nt) {
          acc0 += (double)v0[u];
          acc1 += (double)v1[u];
        }
    }
    if (has0) s_red[sb + lane] = acc0;
    if (has1) s_red[sb + lane + 64] = acc1;
  }
  __syncthreads();

  // 2. Lane r (< D) assembles row r of H and b_r: the re-centred row from the
  //    reduced slots, mapped back with M = [[I, -[c]x], [0, I]] (uncenter_d):
  //    G = H_c M changes columns 3..5, H = M^T G changes rows 3..5 using rows
  //    0..2 of G, which go through LDS.
  const bool row = lane < D;
  const int r = row ? lane : 0;
  double h[D];
  hc_row<CT>(s_red, r, h);
  {
    // h[3..5] -= h[0..2] . column (c - 3) of [cc]x; the zero of each column
    // is left out (adding a product with zero changes no finite value)
    const double h0 = h[0], h1 = h[1], h2 = h[2];
    h[3] -= h1 * cc.z + h2 * (-cc.y);
    h[4] -= h0 * (-cc.z) + h2 * cc.x;
    h[5] -= h0 * cc.y + h1 * (-cc.x);
  }
  if (lane < 3) {
#pragma unroll
    for (int c = 0; c < D; ++c) sG[lane * D + c] = h[c];
  }
  auto bc = [&](int j) { return s_red[j < 6 ? 21 + j : 48 + (j - 6)]; };
  double bv = bc(r);
  __syncthreads();
  if (lane >= 3 && lane < 6) {
    // row (lane - 3) of [cc]x^T is column (lane - 3) of [cc]x: two non-zero
    // entries, on rows i1 < i2 of G
    const int j = lane - 3;
    const int i1 = j == 0 ? 1 : 0, i2 = j == 2 ? 1 : 2;
    const double k1 = j == 0 ? cc.z : j == 1 ? -cc.z : cc.y;
    const double k2 = j == 0 ? -cc.y : j == 1 ? cc.x : -cc.x;
#pragma unroll
    for (int c = 0; c < D; ++c) h[c] -= k1 * sG[i1 * D + c] + k2 * sG[i2 * D + c];
    bv = bv - (k1 * bc(i1) + k2 * bc(i2));
  }

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
    if (row) {
#pragma unroll
      for (int c = 0; c < D; ++c) ps.H[r * D + c] = h[c];
      ps.b[r] = bv;
    }
    if (tid == 0) ps.stats_log[0] = st;
    if constexpr (DUAL) {
      // The dual blocks (CTICP_ H_dual_*, cticp_cpu_impl_.h:202-204) into
      // ps.Hd: H_vel_B 36 | b_vel_B 6 | H_cross_pose_vel_B 36 | H_cross_vel_AB 36.
      // Only the pose rows of H_cross_pose_vel_B carry the re-centring; they
      // map back like b: rows 3..5 -= [c]x^T rows 0..2.
      const double* x = s_red + kCtSlots;
      for (int e = tid; e < 36; e += 64) {
        const int rr = e / 6, cc2 = e % 6;
        const int a0 = rr < cc2 ? rr : cc2, a1 = rr < cc2 ? cc2 : rr;
        ps.Hd[e] = x[a0 * 6 - (a0 * (a0 - 1)) / 2 + (a1 - a0)];
        double cp = x[27 + e];
        if (rr >= 3) {
          const V3d k = skew_col(cc, rr - 3);
          cp -= (k.x * x[27 + cc2] + k.y * x[27 + 6 + cc2]) + k.z * x[27 + 12 + cc2];
        }
        ps.Hd[42 + e] = cp;
        ps.Hd[78 + e] = x[63 + e];
      }
      if (tid < 6) ps.Hd[36 + tid] = x[21 + tid];
    }
    return;
  }

  // 3. pose prior (tree_aligner_base_.h:139-152): lane 0 forms the 6x6 block
  //    and its b, the pose rows add them
  if (has_prior) {
    if (tid == 0) {
      IsoD X = iso_from_34<double>(S.X);
      IsoD ZXi = iso_from_34<double>(S.prior_ZXi);
      double e[6], J[36], OJ[36], Oe[6];
      logmap_d(compose(ZXi, X), e);
      dPoseToPoseError_d(ZXi, X, J);
      for (int rr = 0; rr < 6; ++rr) {
        double se = 0.0;
        for (int k = 0; k < 6; ++k) se += S.prior_info[rr * 6 + k] * e[k];
        Oe[rr] = se;
        for (int c = 0; c < 6; ++c) {
          double sm = 0.0;
          for (int k = 0; k < 6; ++k) sm += S.prior_info[rr * 6 + k] * J[k * 6 + c];
          OJ[rr * 6 + c] = sm;
        }
      }
      for (int rr = 0; rr < 6; ++rr) {
        double sbv = 0.0;
        for (int k = 0; k < 6; ++k) sbv += J[k * 6 + rr] * Oe[k];
        sP[36 + rr] = sbv;
        for (int c = 0; c < 6; ++c) {
          double sm = 0.0;
          for (int k = 0; k < 6; ++k) sm += J[k * 6 + rr] * OJ[k * 6 + c];
          sP[rr * 6 + c] = sm;
        }
      }
    }
    __syncthreads();
    if (lane < 6) {
#pragma unroll
      for (int c = 0; c < 6; ++c) h[c] += sP[lane * 6 + c];
      bv += sP[36 + lane];
    }
  }
  if (row) {
#pragma unroll
    for (int c = 0; c < D; ++c) ps.H[r * D + c] = h[c];
    ps.b[r] = bv;
  }

  // 4. min_good gate (icp_impl_.h:76-80); st is wave-uniform
  if (st.num_inliers + st.num_outliers < sp.min_good_measurements) {
    if (tid == 0) {
      ps.stats_log[it] = st;
      ps.iter = it + 1;
      ps.status = 2;  // failed: compute() returns -it
      if (status_out) status_out[0] = 2;
    }
    for (int m = 1 + tid; m < n_entries; m += 64) {
      members[m].iter = it + 1;
      members[m].status = 2;
    }
    return;
  }

  // 5. damped solve from the rows in registers
  if (!row) {
#pragma unroll
    for (int c = 0; c < D; ++c) h[c] = 0.0;
    bv = 0.0;
  }
  wave_ldlt_solve<D>(h, bv, sp.damping, sp.velocity_damping, lane, sx);
  __syncthreads();

  // 6. Update, termination and the next FP32 cache (updateCache,
  //    icp_impl_.h:16-21), formed in FP64 and rounded once.  Lane 0 forms
  //    exp(dx); lane e < 12 then forms element e (3x4 row-major) of the new
  //    pose and of Xmf = Xf^-1 X with the operation order of compose() and
  //    inverse(); lanes 0..2 form the re-centring anchor.  Xf, and mroot in
  //    the cache, do not change over the alignment and are not rewritten.
  double dx[D];
#pragma unroll
  for (int i = 0; i < D; ++i) dx[i] = sx[i];
  double pn = 0.0, vn = 0.0;
#pragma unroll
  for (int i = 0; i < 6; ++i) pn += dx[i] * dx[i];
  if (CT) {
#pragma unroll
    for (int i = 6; i < D; ++i) vn += dx[i] * dx[i];
  }
  double* sE = sL;        // exp(dx), 3x4
  double* sXc = sL + 12;  // new pose, 3x4
  double* sC = sL + 24;   // anchor, world frame
  if (tid == 0 && !stats_mode) iso_to_34(expmap_d(dx), sE);
  __syncthreads();
  const int ei = lane >> 2, ej = lane & 3;
  float* cache_Xm = reinterpret_cast<float*>(&ps.cache.Xm);
  float* cache_Xmf = reinterpret_cast<float*>(&ps.cache.Xmf);
  const int cache_slot = ej < 3 ? ei * 3 + ej : 9 + ei;  // Iso layout: R row-major, then t
  if (lane < 12) {
    double xc = S.X[lane];
    if (!stats_mode) {
      xc = (sE[ei * 4 + 0] * S.X[0 * 4 + ej] + sE[ei * 4 + 1] * S.X[1 * 4 + ej]) + sE[ei * 4 + 2] * S.X[2 * 4 + ej];
      if (ej == 3) xc = xc + sE[ei * 4 + 3];
      ps.X[lane] = xc;
    }
    sXc[lane] = xc;
    cache_Xm[cache_slot] = (float)xc;
  }
  if (CT && lane < 6) {
    double v = S.v[lane];
    if (!stats_mode) {
      v += sx[6 + lane];
      ps.v[lane] = v;
    }
    ps.cache.v[lane] = (float)v;
  }
  __syncthreads();
  // Xf^-1 = [R^T | -(R^T t)]: entry (i, k) of its rotation is Xf(k, i)
  auto xf_inv = [&](int i, int k) { return S.Xf[k * 4 + i]; };
  auto xf_inv_t = [&](int i) {
    return -((S.Xf[0 * 4 + i] * S.Xf[3] + S.Xf[1 * 4 + i] * S.Xf[7]) + S.Xf[2 * 4 + i] * S.Xf[11]);
  };
  if (lane < 12) {
    double m = (xf_inv(ei, 0) * sXc[0 * 4 + ej] + xf_inv(ei, 1) * sXc[1 * 4 + ej]) + xf_inv(ei, 2) * sXc[2 * 4 + ej];
    if (ej == 3) m = m + xf_inv_t(ei);
    cache_Xmf[cache_slot] = (float)m;
  }
  if (lane < 3) {
    double c = 0.0;
    if (sp.recenter)
      c = ((sXc[lane * 4 + 3] + sXc[lane * 4 + 0] * (double)root_mean.x) + sXc[lane * 4 + 1] * (double)root_mean.y) +
          sXc[lane * 4 + 2] * (double)root_mean.z;
    sC[lane] = c;
    ps.c64[lane] = c;
  }
  __syncthreads();
  if (lane < 3) {
    double cf = 0.0;
    if (sp.recenter)
      cf = ((xf_inv_t(lane) + xf_inv(lane, 0) * sC[0]) + xf_inv(lane, 1) * sC[1]) + xf_inv(lane, 2) * sC[2];
    (&ps.cache.cf.x)[lane] = (float)cf;
  }
  // Forest entries: the same moving pose and velocities, their own Xmf and
  // anchor in their fixed frame.
  for (int m = 1; m < n_entries; ++m) {
    PairState& pm = members[m];
    double* sXfm = sL + 28;  // entry m's Xf, 3x4
    if (lane < 12) sXfm[lane] = pm.Xf[lane];
    __syncthreads();
    auto xfm_inv = [&](int i, int k) { return sXfm[k * 4 + i]; };
    auto xfm_inv_t = [&](int i) {
      return -((sXfm[0 * 4 + i] * sXfm[3] + sXfm[1 * 4 + i] * sXfm[7]) + sXfm[2 * 4 + i] * sXfm[11]);
    };
    float* m_Xm = reinterpret_cast<float*>(&pm.cache.Xm);
    float* m_Xmf = reinterpret_cast<float*>(&pm.cache.Xmf);
    if (lane < 12) {
      double mf = (xfm_inv(ei, 0) * sXc[0 * 4 + ej] + xfm_inv(ei, 1) * sXc[1 * 4 + ej]) + xfm_inv(ei, 2) * sXc[2 * 4 + ej];
      if (ej == 3) mf = mf + xfm_inv_t(ei);
      m_Xmf[cache_slot] = (float)mf;
      m_Xm[cache_slot] = (float)sXc[lane];
    }
    if (CT && lane < 6) pm.cache.v[lane] = (float)(stats_mode ? S.v[lane] : S.v[lane] + sx[6 + lane]);
    if (lane < 3) {
      double cf = 0.0;
      if (sp.recenter) cf = ((xfm_inv_t(lane) + xfm_inv(lane, 0) * sC[0]) + xfm_inv(lane, 1) * sC[1]) + xfm_inv(lane, 2) * sC[2];
      (&pm.cache.cf.x)[lane] = (float)cf;
    }
    __syncthreads();
  }
  if (tid != 0) return;

  st.pert_pose_norm = sqrt(pn);
  st.pert_vel_norm = sqrt(vn);
  ps.stats_log[it] = st;
  ps.iter = it + 1;
  // termination (compute(), tree_aligner_base_.h:100-115): the score
  // criterion (icp_stats_.h:85-102) is asked after every round, the last one
  // included, and a stop it reports wins over the end of the loop (compute()
  // then returns the round index, not max_iterations).  Stats mode runs one
  // round.  The score is FP32 as ICPStats_::score(); the relative change is
  // divided by prev_score + 1e-6 in double, as the reference's double literal
  // promotes it.
  const bool last = stats_mode != 0 || it + 1 >= sp.max_iterations;
  int status = 0;
  if (sp.use_term_crit) {
    const float ni = (float)st.num_inliers;
    const int nl = st.num_inliers + st.num_outliers + st.num_bad + st.num_err_normal + st.num_err_fn;
    const float ratio = nl ? ni / (float)nl : 0.0f;
    const float score = ratio / ((float)st.chi2_kernelized / ni + 1e-6f);
    if (it > 0) {
      const bool pert_ok = (float)st.pert_pose_norm < sp.d_pert;
      const float ds = score - prev_score;
      const bool score_ok = (ds > 0.0f) && (double)ds / ((double)prev_score + 1e-6) < (double)sp.d_score;
      if (pert_ok && score_ok) status = 3;
    }
    ps.prev_score = score;
  }
  if (status == 0 && last) status = 1;
  ps.status = status;
  if (status_out) status_out[0] = status;
  for (int m = 1; m < n_entries; ++m) {
    members[m].iter = it + 1;
    members[m].status = status;
  }
}

// ---- the tile kernel --------------------------------------------------------

// Resident waves per SIMD the tile kernel is compiled for (its register
// budget is 512 / waves VGPRs): 8, the most a SIMD holds, for the rigid form
// with one leaf per lane when its LDS allows 8 (at most 256 staged nodes,
// 64 VGPRs); 4 (128 VGPRs) for the CT and dual forms, for two leaves per lane
// and for deeper staging.
template <bool CT, bool DUAL, int QPL, int LDS_NODES>
constexpr int kTileWaves() {
  return (CT || DUAL || QPL == 2 || LDS_NODES > 256) ? 4 : 8;
}

// Workgroup of the tile kernel: TL / QPL threads, each lane carrying QPL of
// the tile's TL moving leaves (QPL = 2: two descents in flight per lane).
// DUAL (CT only): the dual form of quadraticTermDual, built for bundle
// adjustment; three MFMA products per wave instead of one.
// QPL: moving leaves per lane (1, or 2 for the rigid and CT forms).
// TL: moving leaves per tile, 256 or 128 (one leaf per lane).
template <bool CT, int LDS_NODES, bool DUAL, int QPL, int TL>
__global__ __launch_bounds__(TL / QPL) __attribute__((amdgpu_waves_per_eu(kTileWaves<CT, DUAL, QPL, LDS_NODES>()))) void tiles_kernel(const PairState* __restrict__ pairs,
                                                         const TileDesc* __restrict__ tiles,
                                                         int n_tiles_total,
                                                         float* __restrict__ partials,
                                                         int* __restrict__ corr_out,
                                                         int* __restrict__ depth_out,
                                                         int xcd_remap) {
  static_assert(CT || !DUAL, "the dual form is a CT form");
  static_assert(QPL == 1 || (QPL == 2 && !DUAL && TL == 256), "two leaves per lane: 256-leaf rigid and CT tiles");
  static_assert(TL == 256 || TL == 128, "tiles of 256 or 128 moving leaves");
  constexpr int NT = TL / QPL;         // threads
  constexpr int NWV = NT / 64;         // waves
  constexpr int N = n_slots<CT, DUAL>();
  constexpr int E = CT ? 12 : 6;       // component of sqrt(g) e in w
  constexpr int NW = E + 2;            // components of w
  constexpr int CS = CT ? 90 : 27;     // chi2 slot
  constexpr int HW = (NW + 1) / 2;     // components per transpose half
  constexpr int WS = 68;               // transpose row stride (floats): 16-B aligned, spreads banks
  constexpr int WBUF = DUAL ? 3 * 256 : HW * WS;  // per wave: transpose, then the C tile(s)
  constexpr int kSteps = CT ? 16 : 8;             // MFMA steps per 64 leaves (rigid: two leaf sets per step)
  constexpr int kPacked = CT ? 0 : 8 * 16 + 8;    // rigid: cell offset of the second set's block
  static_assert(HW * WS >= 256 && WBUF >= HW * WS, "the per-wave C tiles reuse the transpose buffer");
  __shared__ float4 s_nodes[LDS_NODES > 0 ? 2 * LDS_NODES : 2];
  __shared__ __align__(16) float s_w[NWV][WBUF];
  __shared__ int s_cnt[NWV][5];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  int cached_pair = -1;
  int n_lds = 0;

  // Tile schedule.  Workgroups are dispatched round-robin over the 8 XCDs of
  // the chip, each with its own L2; with xcd_remap the XCD of workgroup b
  // (b % 8) takes a contiguous eighth of the tile list, so consecutive tiles
  // of one alignment -- which read the same fixed tree -- share an L2.
  // Within its range (the whole list without remap) every workgroup takes
  // one contiguous chunk of tiles: a persistent grid smaller than the tile
  // count then walks consecutive tiles of one alignment and stages the top
  // of that alignment's tree in LDS once, not once per tile.
  const int G = gridDim.x;
  const bool remap = xcd_remap && (G % 8) == 0 && G >= 8;
  const int lo_all = remap ? min((int)(blockIdx.x & 7) * ((n_tiles_total + 7) / 8), n_tiles_total) : 0;
  const int hi_all = remap ? min(lo_all + (n_tiles_total + 7) / 8, n_tiles_total) : n_tiles_total;
  const int n_wg = remap ? G / 8 : G;
  const int j = remap ? (int)(blockIdx.x >> 3) : (int)blockIdx.x;
  const int span = hi_all - lo_all;
  // With at least as many workgroups as tiles (the usual one-per-tile grid)
  // workgroup j takes tile j of the range, without a division; a persistent
  // grid gives workgroup j chunk j of n_wg, sizes differing by at most one
  // (a 64-bit division, once per workgroup of many tiles).
  int c_lo, c_hi;
  if (n_wg >= span) {
    c_lo = lo_all + min(j, span);
    c_hi = lo_all + min(j + 1, span);
  } else {
    c_lo = lo_all + (int)(((long long)span * j) / n_wg);
    c_hi = lo_all + (int)(((long long)span * (j + 1)) / n_wg);
  }
  for (int t = c_lo; t < c_hi; ++t) {
    // Two memory round trips before the descent: the tile's record (its
    // pair, the pair's trees, its first leaf), then the leaf records
    // together with the top-of-tree staging and the pair's status, cache
    // and term parameters.  A finished pair is skipped only after these
    // loads are issued: a skip test right after a status load lets the
    // compiler sink every later load below it, which costs a round trip of
    // its own.  Without a tile list every tile belongs to pair 0.
    const TileDesc d = tiles ? load_tile_desc(tiles + t) : make_tile_desc(pairs[0], 0, t - pairs[0].tile_first, TL);
    const int p = d.pair;
    const PairState& ps = pairs[p];
    glb_f4* nodes = global_f4(d.nodes);
    glb_f4* leaves = global_f4(d.leaves);
    // top-of-tree staging loads: all in flight at once, from clamped
    // addresses (a tree without nodes stages nothing and reads a dummy
    // record), stored after the barrier below.  They are issued for every
    // tile (a later tile of the same pair on a persistent grid re-reads
    // cached lines and stores nothing): loads under a branch would leave the
    // array in scratch memory.
    const bool stage = LDS_NODES > 0 && p != cached_pair;
    constexpr int kStage = LDS_NODES > 0 ? (2 * LDS_NODES + NT - 1) / NT : 1;
    const int n_st = d.n_nodes < LDS_NODES ? d.n_nodes : LDS_NODES;
    glb_v4* st_src = (glb_v4*)(n_st > 0 ? nodes : global_f4(kNoRecord));
    const int st_last = n_st > 0 ? 2 * n_st - 1 : 0;
    v4f st[kStage];
    if constexpr (LDS_NODES > 0) {
#pragma unroll
      for (int k = 0; k < kStage; ++k) st[k] = st_src[min(tid + k * NT, st_last)];
    }
    // lane `tid` carries leaves q[k] = tile base + k * NT + tid
    int q[QPL];
    bool live[QPL];
    float4 la[QPL], lb[QPL];
#pragma unroll
    for (int k = 0; k < QPL; ++k) {
      q[k] = d.leaf0 + k * NT + tid;
      live[k] = q[k] < d.n_leaves && d.n_nodes > 0;
      const int qc = live[k] ? q[k] : 0;  // a tile exists, so the pair has >= 1 leaf
      la[k] = leaves[2 * qc];
      lb[k] = leaves[2 * qc + 1];
    }
    const int status = ps.status;
    const TermCache C = ps.cache;
    const TermParams TP = ps.tp;
    if (stage) {
      // the barrier before the stores keeps an earlier tile's readers of the
      // staged nodes ahead of them; a workgroup's first tile has none
      if (t != c_lo) __syncthreads();
#pragma unroll
      for (int k = 0; k < kStage; ++k)
        if (tid + k * NT < 2 * n_st) ((lds_v4*)s_nodes)[tid + k * NT] = st[k];
      n_lds = n_st;
      cached_pair = p;
      __syncthreads();
    }
    if (status != 0) continue;

    TermRec R[QPL];
    float Jb[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    LeafIn L[QPL];
#pragma unroll
    for (int k = 0; k < QPL; ++k) {
      rec_reject(R[k], -1);  // padding lanes contribute nothing
      L[k].mean = v3(la[k].x, la[k].y, la[k].z);
      L[k].dir = v3(la[k].w, lb[k].x, lb[k].y);
      L[k].dts = lb[k].z;
    }
    glb_f* fdts = global_f(ps.fixed.node_dts);
    auto dts_of = [fdts](int i) { return fdts[i]; };
    auto leaf_terms = [&](const auto& f) {
      if constexpr (QPL == 1) {
        if (live[0]) {
          int depth = 0;
          const int leaf = CT ? ct_term<DUAL>(f, C, TP, L[0], R[0], &depth, Jb, dts_of)
                              : icp_term(f, C, TP, L[0], R[0], &depth);
          if (DUAL && R[0].outcome != kInlier && R[0].outcome != kOutlier)
            for (int i = 0; i < 6; ++i) Jb[i] = 0.0f;
          if (corr_out) corr_out[q[0]] = leaf;
          if (depth_out) depth_out[q[0]] = depth;
        }
      } else {
        // both query points, both descents in lockstep, then both terms
        const V3f zero = v3(0.0f, 0.0f, 0.0f);
        const V3f q0 = live[0] ? (CT ? ct_query(C, L[0]) : icp_query(C, L[0])) : zero;
        const V3f q1 = live[1] ? (CT ? ct_query(C, L[1]) : icp_query(C, L[1])) : zero;
        int leaf0, leaf1, d0, d1;
        float4 a0, b0, a1, b1;
        descend2(f, q0, q1, live[0], live[1], leaf0, leaf1, d0, d1, a0, b0, a1, b1);
        if (live[0]) {
          if constexpr (CT) ct_eval<false>(C, TP, L[0], leaf0, a0, b0, R[0]);
          else icp_eval(C, TP, L[0], q0, leaf0, a0, b0, R[0]);
          if (corr_out) corr_out[q[0]] = leaf0;
          if (depth_out) depth_out[q[0]] = d0;
        }
        if (live[1]) {
          if constexpr (CT) ct_eval<false>(C, TP, L[1], leaf1, a1, b1, R[1]);
          else icp_eval(C, TP, L[1], q1, leaf1, a1, b1, R[1]);
          if (corr_out) corr_out[q[1]] = leaf1;
          if (depth_out) depth_out[q[1]] = d1;
        }
      }
    };
    if (LDS_NODES > 0) {
      // the descent counts its depth only when the launch asks for depths
      if (depth_out && QPL == 1)
        leaf_terms(LdsFetch<true>{nodes, (lds_f4*)s_nodes, n_lds});
      else
        leaf_terms(LdsFetch<false>{nodes, (lds_f4*)s_nodes, n_lds});
    } else {
      leaf_terms(GlobalFetch{nodes});
    }

    // Normal equations of the tile on the matrix cores.  With
    // w = [sqrt(g) J, sqrt(g) Jv (CT), sqrt(g) e, e] per leaf, every
    // accumulator slot is an entry of C = sum_leaves w w^T: the 6x6 (and CT
    // 12x12) blocks of H, b, sum g e^2 (chi2 kernelised) and sum e^2 (chi2).
    // Per wave C is a sum of v_mfma_f32_16x16x4f32 over 64 leaves with the
    // leaves as the K dimension; the leaf vectors are transposed through LDS
    // into the operand layout.  CT (14 components): 16 steps, lane i + 16k
    // holds component i of leaf 16k + m at step m.  Rigid (8 components):
    // the operand's rows 8..15 carry a second set of leaves, so a step takes
    // 8 leaves and a wave 8 steps: lane i + 16k holds component i & 7 of leaf
    // 16k + 8 (i >> 3) + m at step m, and the tile's C is the sum of the two
    // diagonal 8x8 blocks (the off-diagonal blocks mix the two sets and are
    // never read).  The outcome counts come from ballots.
    {
      float* sw = s_w[wave];
      const int ci = lane & 15, cg = lane >> 4;
      floatx4 c0 = {0.0f, 0.0f, 0.0f, 0.0f}, c1 = {0.0f, 0.0f, 0.0f, 0.0f};
      float a[16];
      int cnt[5] = {0, 0, 0, 0, 0};
#pragma unroll
      for (int k = 0; k < QPL; ++k) {
        float w[NW];
        const float sg = R[k].sg;
#pragma unroll
        for (int i = 0; i < 6; ++i) w[i] = sg * R[k].J[i];
        if (CT) {
#pragma unroll
          for (int i = 0; i < 6; ++i) w[6 + i] = sg * R[k].Jv[i];
        }
        w[E] = sg * R[k].e;
        w[E + 1] = R[k].e;
        // The transpose goes through LDS in two halves of HW components, so
        // the buffer stays small enough for four workgroups per CU.
#pragma unroll
        for (int m = 0; m < 16; ++m) a[m] = 0.0f;
#pragma unroll
        for (int half = 0; half < 2; ++half) {
#pragma unroll
          for (int i = 0; i < HW; ++i)
            if (half * HW + i < NW) sw[i * WS + lane] = w[half * HW + i];
          wave_lds_sync();
          const int comp = CT ? ci : (ci & 7);                   // component this lane feeds
          const int first = CT ? cg * 16 : cg * 16 + 8 * (ci >> 3);  // its first leaf
          if (comp >= half * HW && comp < half * HW + HW && comp < NW) {
            const float4* src = reinterpret_cast<const float4*>(sw + (comp - half * HW) * WS + first);
#pragma unroll
            for (int m4 = 0; m4 < kSteps / 4; ++m4) {
              const float4 v = src[m4];
              a[4 * m4 + 0] = v.x;
              a[4 * m4 + 1] = v.y;
              a[4 * m4 + 2] = v.z;
              a[4 * m4 + 3] = v.w;
            }
          }
          wave_lds_sync();
        }
#pragma unroll
        for (int m = 0; m < kSteps; m += 2) {
          c0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[m], a[m], c0, 0, 0, 0);
          c1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[m + 1], a[m + 1], c1, 0, 0, 0);
        }
        cnt[kInlier] += __popcll(__ballot(R[k].outcome == kInlier));
        cnt[kOutlier] += __popcll(__ballot(R[k].outcome == kOutlier));
        cnt[kBad] += __popcll(__ballot(R[k].outcome == kBad));
        cnt[kErrNormal] += __popcll(__ballot(R[k].outcome == kErrNormal));
        cnt[kErrFn] += __popcll(__ballot(R[k].outcome == kErrFn));
      }
      const floatx4 c = c0 + c1;
      if constexpr (DUAL) {
        // z = sqrt(g) J_vel_B: a third transpose, then C_uz = sum w z^T and
        // C_zz = sum z z^T (QPL = 1: a[] still holds this lane's w operands)
        const float sg = R[0].sg;
        float az[16];
#pragma unroll
        for (int m = 0; m < 16; ++m) az[m] = 0.0f;
#pragma unroll
        for (int i = 0; i < 6; ++i) sw[i * WS + lane] = sg * Jb[i];
        wave_lds_sync();
        if (ci < 6) {
          const float4* src = reinterpret_cast<const float4*>(sw + ci * WS + cg * 16);
#pragma unroll
          for (int m4 = 0; m4 < 4; ++m4) {
            const float4 v = src[m4];
            az[4 * m4 + 0] = v.x;
            az[4 * m4 + 1] = v.y;
            az[4 * m4 + 2] = v.z;
            az[4 * m4 + 3] = v.w;
          }
        }
        wave_lds_sync();
        floatx4 cuz = {0.0f, 0.0f, 0.0f, 0.0f}, czz = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int m = 0; m < 16; ++m) {
          cuz = __builtin_amdgcn_mfma_f32_16x16x4f32(a[m], az[m], cuz, 0, 0, 0);
          czz = __builtin_amdgcn_mfma_f32_16x16x4f32(az[m], az[m], czz, 0, 0, 0);
        }
        *reinterpret_cast<float4*>(&sw[256 + ci * 16 + 4 * cg]) = make_float4(cuz[0], cuz[1], cuz[2], cuz[3]);
        *reinterpret_cast<float4*>(&sw[512 + ci * 16 + 4 * cg]) = make_float4(czz[0], czz[1], czz[2], czz[3]);
      }
      // lane (ci, cg) holds C[4 cg + r][ci], stored at ci * 16 + 4 cg + r,
      // in this wave's part of the transpose buffer
      *reinterpret_cast<float4*>(&sw[ci * 16 + 4 * cg]) = make_float4(c[0], c[1], c[2], c[3]);
      if (lane == 0) {
#pragma unroll
        for (int k = 0; k < 5; ++k) s_cnt[wave][k] = cnt[k];
      }
    }
    __syncthreads();
    for (int sl = tid; sl < N; sl += NT) {
      float sv;
      if (sl < CS + 2 || (DUAL && sl >= kCtSlots)) {
        const int cell = (DUAL && sl >= kCtSlots) ? dual_slot_cell(sl) : slot_cell<CT>(sl);
        // a wave's entry: its C cell, plus the second leaf set's for the rigid form
        auto cv = [&](int wv) { return kPacked ? s_w[wv][cell] + s_w[wv][cell + kPacked] : s_w[wv][cell]; };
        if constexpr (NWV == 4) sv = (cv(0) + cv(1)) + (cv(2) + cv(3));
        else if constexpr (NWV == 2) sv = cv(0) + cv(1);
        else sv = cv(0);
      } else {
        const int k = sl - CS - 2;
        if constexpr (NWV == 4) sv = (float)((s_cnt[0][k] + s_cnt[1][k]) + (s_cnt[2][k] + s_cnt[3][k]));
        else if constexpr (NWV == 2) sv = (float)(s_cnt[0][k] + s_cnt[1][k]);
        else sv = (float)s_cnt[0][k];
      }
      partials[(size_t)t * N + sl] = sv;
    }
    // the next tile reuses the transpose buffers: its waves wait here for
    // every reader of this tile's C tiles (no next tile, no wait)
    if (t + 1 < c_hi) __syncthreads();
  }
}

// One workgroup per alignment: FP64 reduction of its partial rows, then
// build-mode publish or solve-mode Gauss-Newton step.
template <bool CT, bool DUAL>
__global__ __launch_bounds__(64) void finish_kernel(PairState* __restrict__ pairs, int pair_first,
                                                          int n_pairs, const float* __restrict__ partials,
                                                          const int* __restrict__ pair_tile_first,
                                                          int* __restrict__ status_out) {
  __shared__ __align__(16) char s_area[FinishLds<CT>::kBytes];
  constexpr int kLoads = (kStateVec4 + 63) / 64;
  __shared__ __align__(16) float4 s_state[64 * kLoads];
  // one workgroup per alignment: the grid is n_pairs wide (no range test,
  // which would split the argument loads into two round trips)
  (void)n_pairs;
  const int p = pair_first + blockIdx.x;
  PairState& ps = pairs[p];
  // One memory round trip for the whole state: every 16-byte load in flight
  // at once, issued together with the tile-range load.  Addresses are
  // clamped and the stores unconditional (the LDS copy is padded to whole
  // waves): a loop, or stores under a condition, would let each load wait
  // for the one before.
  const int tf = pair_tile_first ? pair_tile_first[p] : -1;
  const v4f* src = reinterpret_cast<const v4f*>(&ps);
  v4f sv[kLoads];
#pragma unroll
  for (int k = 0; k < kLoads; ++k) sv[k] = src[min((int)threadIdx.x + 64 * k, kStateVec4 - 1)];
#pragma unroll
  for (int k = 0; k < kLoads; ++k) reinterpret_cast<v4f*>(s_state)[threadIdx.x + 64 * k] = sv[k];
  __syncthreads();
  const PairState& S = *reinterpret_cast<const PairState*>(s_state);
  // forest entries other than the leader are finished by their leader
  if ((S.status != 0) | (S.ct != (CT ? 1 : 0)) | (S.group_n == 0) | (S.group_tiles == 0)) return;
  finish_pair<CT, DUAL>(ps, S, partials, s_area, tf >= 0 ? tf : S.tile_first, status_out ? status_out + p : nullptr);
}

// ---- launch wrappers --------------------------------------------------------

template <bool CT, int LDS, bool DUAL, int QPL, int TL = 256>
static hipError_t launch_t(const LaunchArgs& a, hipStream_t s) {
  if (a.phases & 1) {
    hipLaunchKernelGGL((tiles_kernel<CT, LDS, DUAL, QPL, TL>), dim3(a.grid), dim3(TL / QPL), 0, s, a.pairs,
                       a.tiles, a.n_tiles_total, a.partials, a.corr_out, a.depth_out, a.xcd_remap);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if ((a.phases & 2) && a.n_pairs > 0) {
    hipLaunchKernelGGL((finish_kernel<CT, DUAL>), dim3(a.n_pairs), dim3(64), 0, s, a.pairs, a.pair_first,
                       a.n_pairs, a.partials, a.pair_tile_first, a.status_out);
    return hipGetLastError();
  }
  return hipSuccess;
}

template <bool CT, int LDS>
static hipError_t launch_q(const LaunchArgs& a, hipStream_t s) {
  if (a.tile_leaves == 128) return launch_t<CT, LDS, false, 1, 128>(a, s);
  return a.leaves_per_lane == 2 ? launch_t<CT, LDS, false, 2>(a, s) : launch_t<CT, LDS, false, 1>(a, s);
}

template <int LDS>
static hipError_t launch_dual(const LaunchArgs& a, hipStream_t s) {
  return a.tile_leaves == 128 ? launch_t<true, LDS, true, 1, 128>(a, s) : launch_t<true, LDS, true, 1>(a, s);
}

hipError_t launch_tiles(const LaunchArgs& a, hipStream_t s) {
  if (a.dual) {
    // the dual form caches at most 512 tree nodes in LDS: its third C tile
    // takes the room of the larger caches
    switch (a.lds_nodes) {
      case 0: return launch_dual<0>(a, s);
      case 256: return launch_dual<256>(a, s);
      default: return launch_dual<512>(a, s);
    }
  }
  if (a.ct) {
    switch (a.lds_nodes) {
      case 0: return launch_q<true, 0>(a, s);
      case 256: return launch_q<true, 256>(a, s);
      case 512: return launch_q<true, 512>(a, s);
      default: return launch_q<true, 1024>(a, s);
    }
  }
  switch (a.lds_nodes) {
    case 0: return launch_q<false, 0>(a, s);
    case 256: return launch_q<false, 256>(a, s);
    case 512: return launch_q<false, 512>(a, s);
    case 1024: return launch_q<false, 1024>(a, s);
    default: return launch_q<false, 2048>(a, s);
  }
}

int slots(int ct, int dual) { return dual ? kDualSlots : ct ? kCtSlots : kIcpSlots; }

}  // namespace kdh
