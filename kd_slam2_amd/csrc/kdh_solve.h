// FP64 Gauss-Newton step: quadratic-form assembly, pose prior, damped LDLT
// solve and SE(3) update.  Host and device share this code.
//
// Reference:
//   ICP_::oneRound        kd_slam2/src/lib/kd_slam/icp/icp_impl_.h:53-90
//   CTICP_::oneRound      kd_slam2/src/lib/kd_slam/cticp/cticp_impl_.h:48-92
//   pose prior            icp/tree_aligner_base_.h:139-152
//   expmap / logmap / dlogSO3 / dPoseToPoseError
//                         utils/geometry_3d_impl_.h:48-115, 117-140, 179-194
// The reference solves in FP32 with Eigen's pivoting LDLT; this code solves the
// same damped system in FP64 with an unpivoted LDL^T (the matrix is symmetric
// positive definite once damping > 0 is added), and updates the pose in FP64.
#pragma once

#include "kdh_math.h"

namespace kdh {

// ---- SE(3) in FP64 ----------------------------------------------------------

// Rotation by the axis-angle vector w (Rodrigues).  Below 0.25 rad -- every
// Gauss-Newton step in practice -- R = I + A [w]x + B [w]x^2 with
// A = sin(t)/t and B = (1 - cos(t))/t^2 from their Taylor series in t^2
// (seven terms each; the first omitted ones are below 3e-21) and
// [w]x^2 = w w^T - t^2 I: no square root, division or trigonometric call on
// the solver's serial chain.  Larger angles take the closed form.
KDH_HD M3d rodrigues_d(V3d w) {
  const double t2 = (w.x * w.x + w.y * w.y) + w.z * w.z;
  if (t2 < 0.0625) {
    const double A =
        1.0 + t2 * (-1.0 / 6.0 + t2 * (1.0 / 120.0 + t2 * (-1.0 / 5040.0 + t2 * (1.0 / 362880.0 +
        t2 * (-1.0 / 39916800.0 + t2 * (1.0 / 6227020800.0))))));
    const double B =
        0.5 + t2 * (-1.0 / 24.0 + t2 * (1.0 / 720.0 + t2 * (-1.0 / 40320.0 + t2 * (1.0 / 3628800.0 +
        t2 * (-1.0 / 479001600.0 + t2 * (1.0 / 87178291200.0))))));
    const double d = 1.0 - B * t2;
    const double ax = A * w.x, ay = A * w.y, az = A * w.z;
    const double bx = B * w.x, by = B * w.y;
    M3d R;
    R.m[0] = d + bx * w.x;
    R.m[1] = bx * w.y - az;
    R.m[2] = bx * w.z + ay;
    R.m[3] = bx * w.y + az;
    R.m[4] = d + by * w.y;
    R.m[5] = by * w.z - ax;
    R.m[6] = bx * w.z - ay;
    R.m[7] = by * w.z + ax;
    R.m[8] = d + (B * w.z) * w.z;
    return R;
  }
  double angle = sqrt(t2);
  V3d n = w * (1.0 / angle);
  M3d S = skew(n);
  M3d S2 = mul(S, S);
  double s = sin(angle), c = cos(angle);
  M3d R = m3_identity<double>();
  for (int i = 0; i < 9; ++i) R.m[i] = (R.m[i] + s * S.m[i]) + (1.0 - c) * S2.m[i];
  return R;
}

// expmap: translation = v.head(3), rotation = rodrigues(v.tail(3))
KDH_HD IsoD expmap_d(const double* v) {
  IsoD X;
  X.t = v3(v[0], v[1], v[2]);
  X.R = rodrigues_d(v3(v[3], v[4], v[5]));
  return X;
}

KDH_HD V3d logSO3_d(const M3d& R) {
  double tr = (R.m[0] + R.m[4]) + R.m[8];
  double cos_a = 0.5 * (tr - 1.0);
  cos_a = cos_a < -1.0 ? -1.0 : (cos_a > 1.0 ? 1.0 : cos_a);
  double angle = acos(cos_a);
  V3d ax = v3(R(2, 1) - R(1, 2), R(0, 2) - R(2, 0), R(1, 0) - R(0, 1));
  if (angle < 1e-7) return ax * 0.5;
  return ax * (angle / (2.0 * sin(angle)));
}

KDH_HD void logmap_d(const IsoD& X, double* out6) {
  V3d w = logSO3_d(X.R);
  out6[0] = X.t.x; out6[1] = X.t.y; out6[2] = X.t.z;
  out6[3] = w.x;   out6[4] = w.y;   out6[5] = w.z;
}

// Rotation angle of R in degrees and translation norm, for pose comparisons.
KDH_HD double rot_angle_d(const M3d& R) {
  double tr = (R.m[0] + R.m[4]) + R.m[8];
  double c = 0.5 * (tr - 1.0);
  c = c < -1.0 ? -1.0 : (c > 1.0 ? 1.0 : c);
  return acos(c);
}

// dlogSO3 (3x9, w.r.t. R in column-major order), geometry_3d_impl_.h:57-82.
KDH_HD void dlogSO3_d(const M3d& R, double J[27]) {
  for (int i = 0; i < 27; ++i) J[i] = 0.0;
  V3d ax = v3(R(2, 1) - R(1, 2), R(0, 2) - R(2, 0), R(1, 0) - R(0, 1));
  double tr = (R.m[0] + R.m[4]) + R.m[8];
  double cos_a = 0.5 * (tr - 1.0);
  cos_a = cos_a < -1.0 ? -1.0 : (cos_a > 1.0 ? 1.0 : cos_a);
  double angle = acos(cos_a);
  bool small = angle < 1e-7;
  double f = small ? 0.5 : angle / (2.0 * sin(angle));
  J[0 * 9 + 5] = f;  J[0 * 9 + 7] = -f;
  J[1 * 9 + 6] = f;  J[1 * 9 + 2] = -f;
  J[2 * 9 + 1] = f;  J[2 * 9 + 3] = -f;
  if (!small) {
    double s = sin(angle);
    double c = (s - angle * cos(angle)) / (4.0 * s * s * s);
    for (int k = 0; k < 3; ++k) {
      J[k * 9 + 0] -= c * ax[k];
      J[k * 9 + 4] -= c * ax[k];
      J[k * 9 + 8] -= c * ax[k];
    }
  }
}

// dPoseToPoseError(ZXi, X) (geometry_3d_impl_.h:124-140): 6x6 row-major.
KDH_HD void dPoseToPoseError_d(const IsoD& ZXi, const IsoD& X, double Jout[36]) {
  IsoD P = compose(ZXi, X);
  const M3d& R = ZXi.R;
  // B (12x6): rows 0-2 translation, rows 3+3i.. the i-th column of R_P
  double B[12 * 6];
  for (int i = 0; i < 72; ++i) B[i] = 0.0;
  M3d Rst = mul(R, skew(X.t));
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      B[r * 6 + c] = R(r, c);
      B[r * 6 + 3 + c] = -Rst(r, c);
    }
  for (int i = 0; i < 3; ++i) {
    V3d col = v3(X.R(0, i), X.R(1, i), X.R(2, i));
    M3d Rsc = mul(R, skew(col));
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) B[(3 + 3 * i + r) * 6 + 3 + c] = -Rsc(r, c);
  }
  // dlogmap(P) (6x12) = [I 0; 0 dlogSO3(P.R)]
  double D[27];
  dlogSO3_d(P.R, D);
  for (int r = 0; r < 6; ++r)
    for (int c = 0; c < 6; ++c) {
      double s = 0.0;
      if (r < 3) {
        s = B[r * 6 + c];
      } else {
        for (int k = 0; k < 9; ++k) s += D[(r - 3) * 9 + k] * B[(3 + k) * 6 + c];
      }
      Jout[r * 6 + c] = s;
    }
}

// ---- linear algebra ---------------------------------------------------------

// Solve (A) x = rhs for symmetric positive definite A (n x n, row-major) by an
// unpivoted LDL^T factorisation.  Returns false when a pivot is not positive
// and finite.
template <int N>
KDH_HD bool ldlt_solve_d(const double* A, const double* rhs, double* x) {
  double L[N * N];
  double D[N];
  for (int i = 0; i < N * N; ++i) L[i] = 0.0;
  bool ok = true;
  for (int j = 0; j < N; ++j) {
    double d = A[j * N + j];
    for (int k = 0; k < j; ++k) d -= L[j * N + k] * L[j * N + k] * D[k];
    D[j] = d;
    if (!(d > 0.0) || !(d < 1e300)) ok = false;
    double id = (d != 0.0) ? 1.0 / d : 0.0;
    L[j * N + j] = 1.0;
    for (int i = j + 1; i < N; ++i) {
      double s = A[i * N + j];
      for (int k = 0; k < j; ++k) s -= L[i * N + k] * L[j * N + k] * D[k];
      L[i * N + j] = s * id;
    }
  }
  double y[N];
  for (int i = 0; i < N; ++i) {
    double s = rhs[i];
    for (int k = 0; k < i; ++k) s -= L[i * N + k] * y[k];
    y[i] = s;
  }
  for (int i = 0; i < N; ++i) y[i] = (D[i] != 0.0) ? y[i] / D[i] : 0.0;
  for (int i = N - 1; i >= 0; --i) {
    double s = y[i];
    for (int k = i + 1; k < N; ++k) s -= L[k * N + i] * x[k];
    x[i] = s;
  }
  return ok;
}

// ---- quadratic form assembly ------------------------------------------------

// Unpack an ICPWsItemDiag_ accumulator (21 upper-triangular + 6) into a full
// symmetric block of H (leading dimension ld, offset o) and b (offset o),
// icp_ws_item_.h:36-44.
KDH_HD void unpack_diag_d(const double* acc, double* H, int ld, int o, double* b) {
  int k = 0;
  for (int r = 0; r < 6; ++r)
    for (int c = r; c < 6; ++c, ++k) {
      H[(o + r) * ld + (o + c)] = acc[k];
      H[(o + c) * ld + (o + r)] = acc[k];
    }
  for (int r = 0; r < 6; ++r, ++k) b[o + r] = acc[k];
}

// Map a quadratic form built with the re-centred pose Jacobian
// J_c = [n, (p_m - c) x n_f + (p_f - c) x n_m] back to the reference
// parametrisation: J = J_c M with M = [[I, -[c]x], [0, I]] on the pose block,
// so H = M^T H_c M and b = M^T b_c.  n is the system size (6 or 12); the
// velocity block (rows/cols 6..11) is untouched by M.
template <int N>
KDH_HD void uncenter_d(double* H, double* b, V3d c) {
  // M acts on the pose block only: column block (3..5) of J gets
  // J_t * (-[c]x) added, i.e. for any row vector u: (u M)_{3+j} = u_{3+j} - sum_i u_i [c]x_{ij}.
  M3d Cx = skew(c);
  // H <- H M  (right-multiply: modify columns 3..5)
  for (int r = 0; r < N; ++r) {
    double u0 = H[r * N + 0], u1 = H[r * N + 1], u2 = H[r * N + 2];
    for (int j = 0; j < 3; ++j)
      H[r * N + 3 + j] -= (u0 * Cx(0, j) + u1 * Cx(1, j)) + u2 * Cx(2, j);
  }
  // H <- M^T H (left-multiply: modify rows 3..5)
  for (int cidx = 0; cidx < N; ++cidx) {
    double u0 = H[0 * N + cidx], u1 = H[1 * N + cidx], u2 = H[2 * N + cidx];
    for (int j = 0; j < 3; ++j)
      H[(3 + j) * N + cidx] -= (Cx(0, j) * u0 + Cx(1, j) * u1) + Cx(2, j) * u2;
  }
  // b <- M^T b
  double b0 = b[0], b1 = b[1], b2 = b[2];
  for (int j = 0; j < 3; ++j) b[3 + j] -= (Cx(0, j) * b0 + Cx(1, j) * b1) + Cx(2, j) * b2;
}

}  // namespace kdh
