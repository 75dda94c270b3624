// Single-precision sine and cosine shared, as one source, by the device
// kernels and by the host restatement of the reference aligner.
//
// The continuous-time aligner deskews every moving leaf with R(w*dts)
// (kd_slam2/src/lib/kd_slam/d3/cticp_impl_.h:76-85, Rodrigues in
// utils/geometry_3d_impl_.h:22-46), and the warped mean feeds the KD descent.
// Host libm and the device math library need not agree bit for bit, so both
// sides call this function instead.
//
// Method: the single-precision scheme of S. L. Moshier's Cephes Math Library
// (sinf.c / cosf.c, "Cephes Mathematical Library", 1989-2000): reduction of
// |x| by pi/4 with the three-part constant DP1+DP2+DP3 (Cody-Waite style;
// DP1 has 8 significant bits so j*DP1 is exact for j < 2^16), then the
// published minimax polynomials on [-pi/4, pi/4].  Accuracy is about 1 ulp for
// |x| < 8192; beyond that the reduction loses bits but the result stays a
// deterministic function of x.  Every operation is a single IEEE-754 binary32
// multiply/add/subtract or an exact conversion, evaluated in source order with
// contraction disabled, so host and device produce identical bits.
#pragma once

#include "kdh_math.h"

#pragma clang fp contract(off)

namespace kdh {

KDH_HD void sincos_shared(float x, float* s_out, float* c_out) {
  const float FOPI = 1.27323954473516f;  // 4/pi
  const float DP1 = 0.78515625f;
  const float DP2 = 2.4187564849853515625e-4f;
  const float DP3 = 3.77489497744594108e-8f;
  float ax = x < 0.0f ? -x : x;
  int sign_s = x < 0.0f ? -1 : 1;
  int sign_c = 1;
  // NaN and |x| >= 2^23 skip the reduction (a float-to-integer conversion out
  // of range is undefined); the result is then meaningless but deterministic.
  float q = ax * FOPI;
  unsigned j = (q < 8388608.0f) ? (unsigned)q : 0u;
  float y = (float)j;
  if (j & 1u) {
    j += 1u;
    y += 1.0f;
  }
  j &= 7u;
  if (j > 3u) {
    sign_s = -sign_s;
    sign_c = -sign_c;
    j -= 4u;
  }
  if (j > 1u) sign_c = -sign_c;
  float z = ((ax - y * DP1) - y * DP2) - y * DP3;
  float zz = z * z;
  float ps = (((-1.9515295891e-4f * zz + 8.3321608736e-3f) * zz - 1.6666654611e-1f) * zz) * z + z;
  float pc = ((2.443315711809948e-5f * zz - 1.388731625493765e-3f) * zz + 4.166664568298827e-2f) * zz * zz - 0.5f * zz + 1.0f;
  float s, c;
  if (j == 1u || j == 2u) {
    s = pc;
    c = ps;
  } else {
    s = ps;
    c = pc;
  }
  *s_out = sign_s < 0 ? -s : s;
  *c_out = sign_c < 0 ? -c : c;
}

}  // namespace kdh
