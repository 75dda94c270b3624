// Per-moving-leaf quadratic term of the point-to-plane aligners, for the
// device kernels.
//
// Follows ICP_::quadraticTerm (kd_slam2/src/lib/kd_slam/icp/icp_impl_.h:103-152)
// and CTICP_::quadraticTerm (kd_slam2/src/lib/kd_slam/cticp/cticp_impl_.h:107-181)
// with the 3D traits of d3/icp_.h:41-60 and d3/cticp_impl_.h:11-39, 76-85.
//
// Split of responsibilities:
//  * Everything that decides a discrete outcome -- the leaf warp (CT), the
//    query point p_mif = Xmf*mean, every whichSide test of the descent, the
//    normal-cosine gate, the residual e and its gate, chi2 > kernel_threshold
//    -- is evaluated with the reference formula and operation order, with no
//    contraction, so it matches the host restatement bit for bit.
//  * The Jacobian is re-centred: the rotational column uses (p - c) instead of
//    p, with c a per-iteration centroid chosen by the host, so FP32 terms stay
//    small when the scene is far from the origin.  The host (or the solve
//    stage) maps H_c, b_c back with H = M^T H_c M, b = M^T b_c,
//    M = [[I, -[c]x], [0, I]], in FP64.  c = 0 reproduces the reference
//    Jacobian exactly.
#pragma once

#include "kdh_math.h"
#include "kdh_sincos.h"

#pragma clang fp contract(off)

namespace kdh {

// Accumulator slots per leaf.
//   ICP: 21 upper-triangular J_r J_c g, 6 J_r e g, chi2, chi2*g, 5 one-hot counts
//   CT : 27 pose diag + 27 velocity diag + 36 cross (row-major J_pose_r J_vel_c g)
//        + chi2, chi2*g, 5 counts
enum : int {
  kIcpDiag = 27,
  kIcpSlots = 27 + 2 + 5,      // 34
  kCtDiagPose = 0,
  kCtDiagVel = 27,
  kCtCross = 54,
  kCtSlots = 90 + 2 + 5,       // 97
};

// Outcome of one leaf, as the one-hot counters of ICPStats_ (icp_stats_.h:10-20).
enum : int { kInlier = 0, kOutlier = 1, kBad = 2, kErrNormal = 3, kErrFn = 4 };

struct TermParams {
  float kernel_threshold;
  float max_error;
  float min_normal_cos;
  int stats_mode;
  int recenter;  // 0: the reference's e and J; 1: e and J about the anchor (below)
  float kt_quarter;  // kernel_threshold^(1/4): the outlier weight's square root is kt_quarter / sqrt|e|
};

// Per-iteration inputs of the term (QuadraticTermCache, icp_.h:97-102, plus
// the CT velocity and the re-centring anchor).
//
// Re-centring.  The reference evaluates e = (p_m - p_f).n and
// J_rot = p_m x n_f + p_f x n_m on world points; far from the origin both
// cancel catastrophically in FP32 (a 10 km coordinate carries 1 mm of
// rounding).  With the anchor c = Xm * mroot (mroot = moving tree root mean)
// the same quantities come from frame-local coordinates only:
//   e     = (p_mif - pf) . (Rmf d_m + n_f_local)        (Xf^-1 applied to both)
//   J_rot = Rm (m - mroot) x n_f + Rf (pf - cf) x n_m   (= p_m x n_f + p_f x n_m - c x n)
// with cf = Xf^-1 c.  The FP64 solver maps the form back with the FP64 c.
// The discrete decisions (descent, gates, inlier test) always use the
// reference's own FP32 e.
struct TermCache {
  IsoF Xm;
  IsoF Xf;
  IsoF Xmf;
  float v[6];   // moving velocities [t; w] (CT only)
  V3f mroot;    // anchor in the moving tree frame
  V3f cf;       // anchor in the fixed tree frame
  float vf[6];  // fixed velocities [t; w] (dual form only)
};

// Greedy descent (icp_impl_.h:116-121, node_.h:54-63).  A node fetcher
// (operator()(idx, a, b) loads record idx) also provides descend(q,
// depth_out, a, b): this loop, or one specialised to where its records live,
// with the same decisions in the same order.  Node record layout:
//   a = (mean.x, mean.y, mean.z, dir.x), b = (dir.y, dir.z, left, right)
// with the child links stored as int bits.  Returns the leaf index or -1;
// (a, b) end as the record of the last node fetched, which is the leaf's
// when one is reached (no second fetch of it).
template <typename NodeFetch>
KDH_HD int descend(const NodeFetch& fetch, V3f q, int* depth_out, float4& a, float4& b) {
  int idx = 0;
  int depth = 0;
  fetch(0, a, b);
  for (;;) {
#if defined(__HIP_DEVICE_COMPILE__)
    int l = __float_as_int(b.z);
    int r = __float_as_int(b.w);
#else
    int l, r;
    memcpy(&l, &b.z, 4);
    memcpy(&r, &b.w, 4);
#endif
    if (l < 0 && r < 0) break;
    float dx = q.x - a.x;
    float dy = q.y - a.y;
    float dz = q.z - a.z;
    float s = (a.w * dx + b.x * dy) + b.y * dz;
    idx = s > 0.0f ? l : r;
    ++depth;
    if (idx < 0) break;
    fetch(idx, a, b);
  }
  if (depth_out) *depth_out = depth;
  return idx;
}

// Two greedy descents in lockstep (each exactly `descend`): every step
// issues the next node fetch of both queries before either is needed, so a
// lane keeps two independent gathers in flight.  A query whose `live` is
// false takes no step and returns -1 with depth 0.
template <typename NodeFetch>
KDH_HD void descend2(const NodeFetch& fetch, V3f q0, V3f q1, bool live0, bool live1, int& idx0, int& idx1,
                     int& depth0, int& depth1, float4& a0, float4& b0, float4& a1, float4& b1) {
  idx0 = live0 ? 0 : -1;
  idx1 = live1 ? 0 : -1;
  depth0 = depth1 = 0;
  bool go0 = live0, go1 = live1;
  fetch(0, a0, b0);
  a1 = a0;
  b1 = b0;
  while (go0 || go1) {
    if (go0) {
#if defined(__HIP_DEVICE_COMPILE__)
      const int l = __float_as_int(b0.z), r = __float_as_int(b0.w);
#else
      int l, r;
      memcpy(&l, &b0.z, 4);
      memcpy(&r, &b0.w, 4);
#endif
      if (l < 0 && r < 0) {
        go0 = false;
      } else {
        const float dx = q0.x - a0.x, dy = q0.y - a0.y, dz = q0.z - a0.z;
        const float sd = (a0.w * dx + b0.x * dy) + b0.y * dz;
        idx0 = sd > 0.0f ? l : r;
        ++depth0;
        go0 = idx0 >= 0;
      }
    }
    if (go1) {
#if defined(__HIP_DEVICE_COMPILE__)
      const int l = __float_as_int(b1.z), r = __float_as_int(b1.w);
#else
      int l, r;
      memcpy(&l, &b1.z, 4);
      memcpy(&r, &b1.w, 4);
#endif
      if (l < 0 && r < 0) {
        go1 = false;
      } else {
        const float dx = q1.x - a1.x, dy = q1.y - a1.y, dz = q1.z - a1.z;
        const float sd = (a1.w * dx + b1.x * dy) + b1.y * dz;
        idx1 = sd > 0.0f ? l : r;
        ++depth1;
        go1 = idx1 >= 0;
      }
    }
    if (go0) fetch(idx0, a0, b0);
    if (go1) fetch(idx1, a1, b1);
  }
}

// rodrigues(w) (geometry_3d_impl_.h:22-46): identity below 1e-9, otherwise
// R = I + sin(a) S + (1 - cos(a)) S^2 with S = skew(w / |w|).
struct Rodrigues {
  M3f R;
  V3f n;
  float angle;
  bool small;
};

// Below 1e-9 rad: R = I, S = 0, n = 0, small = true.  Written without a
// branch (the general form on a safe angle, then a per-element choice of
// values): a branch that assigns the matrices on two paths leaves them in
// private scratch memory on the device.
KDH_HD Rodrigues rodrigues_f(V3f w) {
  Rodrigues o;
  o.angle = sqrtf(sqnorm(w));
  const bool small = o.angle < 1e-9f;
  o.small = small;
  const float inv = 1.0f / (small ? 1.0f : o.angle);
  const V3f n = w * inv;
  float s, c;
  sincos_shared(o.angle, &s, &c);
  const float omc = 1.0f - c;
  // (I + s S) + (1 - c) S^2 entry by entry, without the products with an
  // exact zero of I or S that the 3x3 forms carry: each entry of S S is then
  // one product n_i n_j off the diagonal and -(n_j^2 + n_k^2) on it, and
  // I + s S adds exact zeros.  Leaving out an exact zero changes no value
  // but the sign of a zero result, which no comparison, and no sum with a
  // nonzero term, can see: the warp keeps its bits.
  const float xx = n.x * n.x, yy = n.y * n.y, zz = n.z * n.z;
  const float xy = n.x * n.y, xz = n.x * n.z, yz = n.y * n.z;
  const float r[9] = {1.0f + omc * -(zz + yy), omc * xy - s * n.z,      s * n.y + omc * xz,
                      s * n.z + omc * xy,      1.0f + omc * -(zz + xx), omc * yz - s * n.x,
                      omc * xz - s * n.y,      s * n.x + omc * yz,      1.0f + omc * -(yy + xx)};
#pragma unroll
  for (int i = 0; i < 9; ++i) o.R.m[i] = small ? ((i & 3) == 0 ? 1.0f : 0.0f) : r[i];
  o.n = small ? v3(0.0f, 0.0f, 0.0f) : n;
  return o;
}

// Tolerance-path forms of t0 + R p, R^T v, a x b and a . b whose
// multiply-adds fuse: for the Jacobians and the re-centred residual, which no
// decision reads.  The decision values use the unfused helpers in the
// reference's order.  Contraction is per source expression (contract(on), not
// fast), so an expression fuses the same way in every kernel it is inlined
// into: the CT and dual builds keep equal bits on their shared terms.
KDH_HD V3f fused_from(const M3f& R, V3f p, V3f t0) {
#pragma clang fp contract(on)
  return v3(t0.x + R.m[0] * p.x + R.m[1] * p.y + R.m[2] * p.z, t0.y + R.m[3] * p.x + R.m[4] * p.y + R.m[5] * p.z,
            t0.z + R.m[6] * p.x + R.m[7] * p.y + R.m[8] * p.z);
}
KDH_HD V3f fused_mul_t(const M3f& R, V3f v) {
#pragma clang fp contract(on)
  return v3(R.m[0] * v.x + R.m[3] * v.y + R.m[6] * v.z, R.m[1] * v.x + R.m[4] * v.y + R.m[7] * v.z,
            R.m[2] * v.x + R.m[5] * v.y + R.m[8] * v.z);
}
KDH_HD V3f fused_cross(V3f a, V3f b) {
#pragma clang fp contract(on)
  return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
KDH_HD float fused_dot(V3f a, V3f b) {
#pragma clang fp contract(on)
  return a.x * b.x + a.y * b.y + a.z * b.z;
}

// Rotational velocity Jacobian (tolerance path: no decision reads it, so
// multiply-adds may fuse).  d(R(w) u)/dw = -R [u]x B with
// B = n n^T + (R^T - I) [n]x / angle (Gallego & Yezzi; the reference's dRxp,
// geometry_3d_impl_.h:143-174), and B = I below 1e-9 rad, where R = I.  The
// kernels need it only transposed and applied to a vector,
// (d(R u)/dw)^T v = B^T (u x R^T v), two such products with one rotation at a
// time, so this returns
//   B^T (u1 x R^T v1 + u2 x R^T v2),
// with B^T s = n (n . s) - n x (R s - s) / angle: no 3x3 product is formed.
KDH_HD V3f vel_rot_jac(const Rodrigues& rd, V3f u1, V3f v1, V3f u2, V3f v2) {
#pragma clang fp contract(on)
  const float* R = rd.R.m;
  const float q1x = R[0] * v1.x + R[3] * v1.y + R[6] * v1.z;
  const float q1y = R[1] * v1.x + R[4] * v1.y + R[7] * v1.z;
  const float q1z = R[2] * v1.x + R[5] * v1.y + R[8] * v1.z;
  const float q2x = R[0] * v2.x + R[3] * v2.y + R[6] * v2.z;
  const float q2y = R[1] * v2.x + R[4] * v2.y + R[7] * v2.z;
  const float q2z = R[2] * v2.x + R[5] * v2.y + R[8] * v2.z;
  const float sx = (u1.y * q1z - u1.z * q1y) + (u2.y * q2z - u2.z * q2y);
  const float sy = (u1.z * q1x - u1.x * q1z) + (u2.z * q2x - u2.x * q2z);
  const float sz = (u1.x * q1y - u1.y * q1x) + (u2.x * q2y - u2.y * q2x);
  // d = R s - s
  const float dx = R[0] * sx + R[1] * sy + R[2] * sz - sx;
  const float dy = R[3] * sx + R[4] * sy + R[5] * sz - sy;
  const float dz = R[6] * sx + R[7] * sy + R[8] * sz - sz;
  const V3f n = rd.n;
  const float ia = 1.0f / (rd.small ? 1.0f : rd.angle);
  const float ns = n.x * sx + n.y * sy + n.z * sz;
  const float bx = n.x * ns - (n.y * dz - n.z * dy) * ia;
  const float by = n.y * ns - (n.z * dx - n.x * dz) * ia;
  const float bz = n.z * ns - (n.x * dy - n.y * dx) * ia;
  return rd.small ? v3(sx, sy, sz) : v3(bx, by, bz);
}

}  // namespace kdh

namespace kdh {

// One moving leaf as the kernels see it: (mean, dir) in the moving tree frame
// and, for CT, the leaf's mean stamp offset narrowed to float exactly as the
// reference's `const Scalar dt = moving_leaf._dts_mean` (cticp_impl_.h:158).
struct LeafIn {
  V3f mean;
  V3f dir;
  float dts;
};

// Compact result of one leaf: everything the accumulator slots are formed
// from.  The robust weight gamma travels as its square root sg: every slot is
// a product of two entries of w = [sg J, sg Jv, sg e, e], so the kernels never
// need gamma itself.  A rejected leaf has sg = 0, e = 0, J = 0 and its outcome
// set, so every product slot expands to 0 and only its one-hot counter is 1.
struct TermRec {
  float J[6];   // pose Jacobian (re-centred rotational part)
  float Jv[6];  // velocity Jacobian (CT)
  float e;
  float sg;     // sqrt(gamma)
  int outcome;  // kInlier .. kErrFn
};

KDH_HD void rec_reject(TermRec& r, int outcome) {
  for (int i = 0; i < 6; ++i) {
    r.J[i] = 0.0f;
    r.Jv[i] = 0.0f;
  }
  r.e = 0.0f;
  r.sg = 0.0f;
  r.outcome = outcome;
}

// Packing of the 6x6 upper triangles in ICPWsItemDiag_::pack (row by row;
// icp_ws_item_.h:26-34): row and column of the k-th entry.
KDH_HD int tri_r(int k) {
  return k < 6 ? 0 : k < 11 ? 1 : k < 15 ? 2 : k < 18 ? 3 : k < 20 ? 4 : 5;
}
KDH_HD int tri_c(int k) {
  const int r = tri_r(k);
  const int start = r == 0 ? 0 : r == 1 ? 6 : r == 2 ? 11 : r == 3 ? 15 : r == 4 ? 18 : 20;
  return r + (k - start);
}

// Robust weight and outcome (icp_impl_.h:138-151): the inlier test on the
// reference's FP32 chi2 = e_ref^2 decides the outcome bit for bit; the weight
// gamma = sqrt(kernel_threshold / e^2) of an outlier, with e the residual
// R.e that enters b (equal to e_ref unless re-centred), is kept as its square
// root kt_quarter / sqrt|e|, from the hardware reciprocal square root (a few
// ulp from the reference's correctly rounded division and square root: H, b
// and chi2 are tolerance results, no decision reads the weight).
// The dual form weighs differently (quadraticTermDual, cticp_impl_.h:240-248):
// outliers are always weighted, inliers drop to 0 in stats mode.
KDH_HD float rsqrt_f(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rsqf(x);
#else
  return 1.0f / sqrtf(x);
#endif
}
template <bool DUAL = false>
KDH_HD void classify(TermRec& R, float e_ref, const TermParams& P) {
  if (e_ref * e_ref > P.kernel_threshold) {
    R.outcome = kOutlier;
    R.sg = (P.stats_mode && !DUAL) ? 0.0f : P.kt_quarter * rsqrt_f(fabsf(R.e));
  } else {
    R.outcome = kInlier;
    R.sg = (P.stats_mode && DUAL) ? 0.0f : 1.0f;
  }
}

// Rigid ICP term (icp_impl_.h:103-152), in three steps: the query point
// (icp_query), the descent, and the term at the fixed leaf the descent
// reached (icp_eval: leaf index and its record (a, b), or -1).
KDH_HD V3f icp_query(const TermCache& C, const LeafIn& L) { return apply_iso(C.Xmf, L.mean); }

KDH_HD void icp_eval(const TermCache& C, const TermParams& P, const LeafIn& L, V3f p_mif, int leaf, float4 a,
                     float4 b, TermRec& R) {
  if (leaf < 0) {
    rec_reject(R, kBad);
    return;
  }
  const V3f pf = v3(a.x, a.y, a.z);
  const V3f nf = v3(a.w, b.x, b.y);
  const V3f p_m = apply_iso(C.Xm, L.mean);
  const V3f n_m = apply_rot(C.Xm, L.dir);
  const V3f p_f = apply_iso(C.Xf, pf);
  const V3f n_f = apply_rot(C.Xf, nf);
  // errorAndJacobian (d3/icp_.h:41-60)
  if (dot(n_f, n_m) < P.min_normal_cos) {
    rec_reject(R, kErrNormal);
    return;
  }
  const V3f n = n_m + n_f;
  const float e = dot(p_m - p_f, n);
  if (fabsf(e) > P.max_error) {
    rec_reject(R, kErrFn);
    return;
  }
  // Reference mode has mroot = cf = 0 and starts from the full translations,
  // which makes jr p_m x n_f + p_f x n_m (up to the fused roundings).
  const V3f z = v3(0.0f, 0.0f, 0.0f);
  const V3f jr = fused_cross(fused_from(C.Xm.R, L.mean - C.mroot, P.recenter ? z : C.Xm.t), n_f) +
                 fused_cross(fused_from(C.Xf.R, pf - C.cf, P.recenter ? z : C.Xf.t), n_m);
  const float e_loc = fused_dot(p_mif - pf, fused_from(C.Xmf.R, L.dir, nf));
  R.e = P.recenter ? e_loc : e;
  R.J[0] = n.x; R.J[1] = n.y; R.J[2] = n.z;
  R.J[3] = jr.x; R.J[4] = jr.y; R.J[5] = jr.z;
  for (int i = 0; i < 6; ++i) R.Jv[i] = 0.0f;
  classify<false>(R, e, P);
}

// The whole rigid term for one leaf.  Returns the fixed leaf index reached by
// the descent (-1 for num_bad).
template <typename NodeFetch>
KDH_HD int icp_term(const NodeFetch& fetch, const TermCache& C, const TermParams& P, const LeafIn& L, TermRec& R,
                    int* depth_out) {
  const V3f p_mif = icp_query(C, L);
  float4 a, b;
  const int leaf = fetch.descend(p_mif, depth_out, a, b);
  icp_eval(C, P, L, p_mif, leaf, a, b, R);
  return leaf;
}

// Continuous-time term (cticp_impl_.h:107-181).  With DUAL also the
// quadraticTermDual extra (cticp_impl_.h:185-293): Jb, the Jacobian with
// respect to the fixed side's velocities C.vf at the fixed leaf's own stamp
// (dts_of(leaf)), errorAndJacobianVelocityB (d3/cticp_impl_.h:43-73):
//   J_t = -dt n^T Rb ; J_w = dt (-n^T Rb dR_pb + dp^T Rb dR_nb)
// with the fixed leaf warped by vf; re-centred, Rb^T n = Rmf wdir + nb_moved
// and Rb^T dp = p_mif - pb_moved, in the fixed tree's frame.  In three steps
// like the rigid term: query point, descent, term at the fixed leaf.

// The CT query point: the leaf warped by applyVelocities
// (d3/cticp_impl_.h:76-85), then moved into the fixed tree's frame.
KDH_HD void ct_warp(const TermCache& C, const LeafIn& L, Rodrigues& rd, V3f& wmean, V3f& wdir) {
  const float dt = L.dts;
  const V3f t_delta = v3(C.v[0] * dt, C.v[1] * dt, C.v[2] * dt);
  const V3f r_delta = v3(C.v[3] * dt, C.v[4] * dt, C.v[5] * dt);
  rd = rodrigues_f(r_delta);
  wmean = mul(rd.R, L.mean) + t_delta;
  wdir = mul(rd.R, L.dir);
}
KDH_HD V3f ct_query(const TermCache& C, const LeafIn& L) {
  Rodrigues rd;
  V3f wmean, wdir;
  ct_warp(C, L, rd, wmean, wdir);
  return apply_iso(C.Xmf, wmean);
}

// The CT term at the fixed leaf the descent reached (leaf index and its
// record (a, b), or -1), given the leaf's warp (rd, wmean, wdir) and query
// point p_mif.
template <bool DUAL = false, typename DtsFetch = int>
KDH_HD void ct_eval_warped(const TermCache& C, const TermParams& P, const LeafIn& L, const Rodrigues& rd,
                           V3f wmean, V3f wdir, V3f p_mif, int leaf, float4 a, float4 b, TermRec& R,
                           float* Jb = nullptr, const DtsFetch& dts_of = 0) {
  if (leaf < 0) {
    rec_reject(R, kBad);
    return;
  }
  const float dt = L.dts;
  const V3f pf = v3(a.x, a.y, a.z);
  const V3f nf = v3(a.w, b.x, b.y);
  const V3f p_m = apply_iso(C.Xm, wmean);
  const V3f n_m = apply_rot(C.Xm, wdir);
  const V3f p_f = apply_iso(C.Xf, pf);
  const V3f n_f = apply_rot(C.Xf, nf);
  if (dot(n_f, n_m) < P.min_normal_cos) {
    rec_reject(R, kErrNormal);
    return;
  }
  const V3f n = n_m + n_f;
  const V3f dp = p_m - p_f;
  const float e = dot(dp, n);
  if (fabsf(e) > P.max_error) {
    rec_reject(R, kErrFn);
    return;
  }
  // As in icp_term; re-centred, Ra^T dp = Rmf^T (p_mif - pf) as well.
  const M3f& Ra = C.Xm.R;
  const V3f z = v3(0.0f, 0.0f, 0.0f);
  const V3f jr = fused_cross(fused_from(C.Xm.R, wmean - C.mroot, P.recenter ? z : C.Xm.t), n_f) +
                 fused_cross(fused_from(C.Xf.R, pf - C.cf, P.recenter ? z : C.Xf.t), n_m);
  const V3f dl = p_mif - pf;
  const float e_loc = fused_dot(dl, fused_from(C.Xmf.R, wdir, nf));
  R.e = P.recenter ? e_loc : e;
  const V3f dpR = P.recenter ? fused_mul_t(C.Xmf.R, dl) : fused_mul_t(Ra, dp);
  R.J[0] = n.x; R.J[1] = n.y; R.J[2] = n.z;
  R.J[3] = jr.x; R.J[4] = jr.y; R.J[5] = jr.z;
  // errorAndJacobianVelocityA (d3/cticp_impl_.h:11-39) at the unwarped leaf:
  //   J_t = dt n^T Ra ; J_w = dt (n^T Ra dR_p + dp^T Ra dR_n)
  const V3f nR = fused_mul_t(Ra, n);
  const V3f jw = vel_rot_jac(rd, L.mean, nR, L.dir, dpR);
  R.Jv[0] = nR.x * dt; R.Jv[1] = nR.y * dt; R.Jv[2] = nR.z * dt;
  R.Jv[3] = jw.x * dt; R.Jv[4] = jw.y * dt; R.Jv[5] = jw.z * dt;
  if constexpr (DUAL) {
    const float dtb = dts_of(leaf);
    const V3f tb = v3(C.vf[0] * dtb, C.vf[1] * dtb, C.vf[2] * dtb);
    const Rodrigues rb = rodrigues_f(v3(C.vf[3] * dtb, C.vf[4] * dtb, C.vf[5] * dtb));
    const V3f pb_moved = mul(rb.R, pf) + tb;
    const V3f nb_moved = mul(rb.R, nf);
    V3f nbR, dpbR;  // Rb^T n and Rb^T dp
    if (P.recenter) {
      nbR = apply_rot(C.Xmf, wdir) + nb_moved;
      dpbR = p_mif - pb_moved;
    } else {
      const M3f& Rb = C.Xf.R;
      const V3f nb = n_m + apply_rot(C.Xf, nb_moved);
      const V3f dpb = p_m - apply_iso(C.Xf, pb_moved);
      nbR = mul_t(Rb, nb);
      dpbR = mul_t(Rb, dpb);
    }
    const V3f jwb = vel_rot_jac(rb, nf, dpbR, pf, -nbR);
    Jb[0] = -nbR.x * dtb; Jb[1] = -nbR.y * dtb; Jb[2] = -nbR.z * dtb;
    Jb[3] = jwb.x * dtb; Jb[4] = jwb.y * dtb; Jb[5] = jwb.z * dtb;
  }
  classify<DUAL>(R, e, P);
}

// The same, forming the warp again from the leaf (the same operations on the
// same values, so the same bits as the query's): the two-leaves-per-lane
// path keeps no warp across its descents.
template <bool DUAL = false, typename DtsFetch = int>
KDH_HD void ct_eval(const TermCache& C, const TermParams& P, const LeafIn& L, int leaf, float4 a, float4 b,
                    TermRec& R, float* Jb = nullptr, const DtsFetch& dts_of = 0) {
  Rodrigues rd;
  V3f wmean, wdir;
  ct_warp(C, L, rd, wmean, wdir);
  ct_eval_warped<DUAL>(C, P, L, rd, wmean, wdir, apply_iso(C.Xmf, wmean), leaf, a, b, R, Jb, dts_of);
}

// The whole CT term for one leaf (the warp formed once).  Returns the fixed
// leaf index reached by the descent (-1 for num_bad).
template <bool DUAL = false, typename NodeFetch, typename DtsFetch = int>
KDH_HD int ct_term(const NodeFetch& fetch, const TermCache& C, const TermParams& P, const LeafIn& L, TermRec& R,
                   int* depth_out, float* Jb = nullptr, const DtsFetch& dts_of = 0) {
  Rodrigues rd;
  V3f wmean, wdir;
  ct_warp(C, L, rd, wmean, wdir);
  const V3f p_mif = apply_iso(C.Xmf, wmean);
  float4 a, b;
  const int leaf = fetch.descend(p_mif, depth_out, a, b);
  ct_eval_warped<DUAL>(C, P, L, rd, wmean, wdir, p_mif, leaf, a, b, R, Jb, dts_of);
  return leaf;
}

}  // namespace kdh
