// Small fixed-size vector / matrix / isometry types for host and device code.
//
// Every operation is written with an explicit evaluation order, because the
// discrete decisions of the aligner (KD descent side, normal-cosine gate,
// residual gate, inlier test) must come out bit-identical to a host
// restatement of the reference CPU aligner.  The order followed is the one the
// reference's Eigen expressions evaluate to for 3-vectors and 3x3 matrices:
//   dot(a,b)        = (a0*b0 + a1*b1) + a2*b2
//   (M*v)_r         = (M_r0*v0 + M_r1*v1) + M_r2*v2
//   applyIsometry_  = ((t_r + M_r0*p0) + M_r1*p1) + M_r2*p2
//                     (kd_slam2/src/lib/kd_slam/utils/geometry_.h:11-25)
//   applyRotation_  = ((0 + M_r0*p0) + M_r1*p1) + M_r2*p2
// All translation units that include this header are compiled with
// -ffp-contract=off, so no multiply-add is fused behind our back.
#pragma once

#include <stdint.h>
#include <math.h>

#if defined(__HIP__)
#include <hip/hip_runtime.h>
#define KDH_HD __host__ __device__ __forceinline__
#else
#define KDH_HD inline
#include <string.h>
struct float4 {
  float x, y, z, w;
};
#endif

#pragma clang fp contract(off)

namespace kdh {

template <typename T>
struct V3 {
  T x, y, z;
  KDH_HD T operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};

using V3f = V3<float>;
using V3d = V3<double>;

template <typename T> KDH_HD V3<T> v3(T x, T y, T z) { V3<T> r; r.x = x; r.y = y; r.z = z; return r; }
template <typename T> KDH_HD V3<T> operator+(V3<T> a, V3<T> b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
template <typename T> KDH_HD V3<T> operator-(V3<T> a, V3<T> b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
template <typename T> KDH_HD V3<T> operator*(V3<T> a, T s) { return v3(a.x * s, a.y * s, a.z * s); }
template <typename T> KDH_HD V3<T> operator-(V3<T> a) { return v3(-a.x, -a.y, -a.z); }
template <typename T> KDH_HD T dot(V3<T> a, V3<T> b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
template <typename T> KDH_HD T sqnorm(V3<T> a) { return (a.x * a.x + a.y * a.y) + a.z * a.z; }
// Eigen cross(): (a1 b2 - a2 b1, a2 b0 - a0 b2, a0 b1 - a1 b0)
template <typename T> KDH_HD V3<T> cross(V3<T> a, V3<T> b) {
  return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// Row-major 3x3.
template <typename T>
struct M3 {
  T m[9];
  KDH_HD T operator()(int r, int c) const { return m[r * 3 + c]; }
  KDH_HD T& operator()(int r, int c) { return m[r * 3 + c]; }
};
using M3f = M3<float>;
using M3d = M3<double>;

template <typename T> KDH_HD M3<T> m3_identity() {
  M3<T> r;
  for (int i = 0; i < 9; ++i) r.m[i] = T(0);
  r.m[0] = r.m[4] = r.m[8] = T(1);
  return r;
}
template <typename T> KDH_HD M3<T> m3_zero() {
  M3<T> r;
  for (int i = 0; i < 9; ++i) r.m[i] = T(0);
  return r;
}
template <typename T> KDH_HD V3<T> mul(const M3<T>& M, V3<T> v) {
  return v3((M.m[0] * v.x + M.m[1] * v.y) + M.m[2] * v.z,
            (M.m[3] * v.x + M.m[4] * v.y) + M.m[5] * v.z,
            (M.m[6] * v.x + M.m[7] * v.y) + M.m[8] * v.z);
}
// M^T v
template <typename T> KDH_HD V3<T> mul_t(const M3<T>& M, V3<T> v) {
  return v3((M.m[0] * v.x + M.m[3] * v.y) + M.m[6] * v.z,
            (M.m[1] * v.x + M.m[4] * v.y) + M.m[7] * v.z,
            (M.m[2] * v.x + M.m[5] * v.y) + M.m[8] * v.z);
}
template <typename T> KDH_HD M3<T> mul(const M3<T>& A, const M3<T>& B) {
  M3<T> R;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c)
      R.m[r * 3 + c] = (A.m[r * 3 + 0] * B.m[0 * 3 + c] + A.m[r * 3 + 1] * B.m[1 * 3 + c]) + A.m[r * 3 + 2] * B.m[2 * 3 + c];
  return R;
}
template <typename T> KDH_HD M3<T> transpose(const M3<T>& A) {
  M3<T> R;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) R.m[r * 3 + c] = A.m[c * 3 + r];
  return R;
}
template <typename T> KDH_HD M3<T> skew(V3<T> v) {
  M3<T> m;
  m.m[0] = T(0);  m.m[1] = -v.z; m.m[2] = v.y;
  m.m[3] = v.z;   m.m[4] = T(0); m.m[5] = -v.x;
  m.m[6] = -v.y;  m.m[7] = v.x;  m.m[8] = T(0);
  return m;
}

// Rigid transform: rotation R (row-major) and translation t.  A 3x4 row-major
// float array [R | t] is the wire form used across the C ABI.
template <typename T>
struct Iso {
  M3<T> R;
  V3<T> t;
};
using IsoF = Iso<float>;
using IsoD = Iso<double>;

template <typename T> KDH_HD Iso<T> iso_identity() {
  Iso<T> X;
  X.R = m3_identity<T>();
  X.t = v3(T(0), T(0), T(0));
  return X;
}

// applyIsometry_ (geometry_.h:11-25): start from t, add M(r,c)*p(c) for c=0..2.
template <typename T> KDH_HD V3<T> apply_iso(const Iso<T>& X, V3<T> p) {
  return v3(((X.t.x + X.R.m[0] * p.x) + X.R.m[1] * p.y) + X.R.m[2] * p.z,
            ((X.t.y + X.R.m[3] * p.x) + X.R.m[4] * p.y) + X.R.m[5] * p.z,
            ((X.t.z + X.R.m[6] * p.x) + X.R.m[7] * p.y) + X.R.m[8] * p.z);
}
// applyRotation_ (geometry_.h:28-37): same loop starting from zero.
template <typename T> KDH_HD V3<T> apply_rot(const Iso<T>& X, V3<T> p) {
  return v3(((T(0) + X.R.m[0] * p.x) + X.R.m[1] * p.y) + X.R.m[2] * p.z,
            ((T(0) + X.R.m[3] * p.x) + X.R.m[4] * p.y) + X.R.m[5] * p.z,
            ((T(0) + X.R.m[6] * p.x) + X.R.m[7] * p.y) + X.R.m[8] * p.z);
}
// The same loop starting from a given vector: apply_from(X, p, X.t) is
// apply_iso(X, p) and apply_from(X, p, 0) is apply_rot(X, p), bit for bit.
template <typename T> KDH_HD V3<T> apply_from(const Iso<T>& X, V3<T> p, V3<T> t0) {
  return v3(((t0.x + X.R.m[0] * p.x) + X.R.m[1] * p.y) + X.R.m[2] * p.z,
            ((t0.y + X.R.m[3] * p.x) + X.R.m[4] * p.y) + X.R.m[5] * p.z,
            ((t0.z + X.R.m[6] * p.x) + X.R.m[7] * p.y) + X.R.m[8] * p.z);
}
template <typename T> KDH_HD Iso<T> compose(const Iso<T>& A, const Iso<T>& B) {
  Iso<T> C;
  C.R = mul(A.R, B.R);
  C.t = mul(A.R, B.t) + A.t;
  return C;
}
template <typename T> KDH_HD Iso<T> inverse(const Iso<T>& A) {
  Iso<T> C;
  C.R = transpose(A.R);
  C.t = -mul(C.R, A.t);
  return C;
}
template <typename To, typename From> KDH_HD Iso<To> iso_cast(const Iso<From>& A) {
  Iso<To> C;
  for (int i = 0; i < 9; ++i) C.R.m[i] = To(A.R.m[i]);
  C.t = v3(To(A.t.x), To(A.t.y), To(A.t.z));
  return C;
}
// 3x4 row-major [R|t] <-> Iso
template <typename T, typename S> KDH_HD Iso<T> iso_from_34(const S* a) {
  Iso<T> X;
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) X.R.m[r * 3 + c] = T(a[r * 4 + c]);
  }
  X.t = v3(T(a[3]), T(a[7]), T(a[11]));
  return X;
}
template <typename T, typename S> KDH_HD void iso_to_34(const Iso<T>& X, S* a) {
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) a[r * 4 + c] = S(X.R.m[r * 3 + c]);
  }
  a[3] = S(X.t.x); a[7] = S(X.t.y); a[11] = S(X.t.z);
}

}  // namespace kdh
