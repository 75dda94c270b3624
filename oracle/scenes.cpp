// ORACLE -- test infrastructure only.  Nothing in the product links or calls
// this code: only tests/ load it.
//
// Restatements of the scene and scanner generators that the reference's own
// end-to-end tests use to state their known answers:
//   generateScene / addItem       kd_slam2/src/lib/kd_slam/d3/typedefs.cpp:9-34, 37-64
//   Scanner::makeScan             kd_slam2/src/lib/kd_slam/d3/scanner.cpp:5-83
//   ScannerParams (defaults)      kd_slam2/src/lib/kd_slam/d3/scanner.h:6-16
// used by apps/tests/icp/kd_icp_main.cpp:36-64 (rigid recovery of
// x_gt = (1, 2, 1, 0.1, 0.3, 0.1)) and apps/tests/cticp/kd_cticp_main.cpp:50-82
// (CT recovery of X_gt = expmap(0.1, 2, 0, 0, 0, 0.4), v_gt = (1, 0, 0, 0, 0, 2)).
//
// Random numbers.  The reference draws Eigen's Random() (Eigen 3.4:
// x + (y - x) * float(rand()) / float(RAND_MAX) on [-1, 1], coefficient by
// coefficient) and drand48(), both without seeding, so a fresh process starts
// from glibc's unseeded states: rand() as after srand(1), and drand48() from a
// zero state (seed48 of three zero words).  Both are reset here before every
// scene, so the scene is the one a fresh run of the reference test sees.
// Rotations from an angle and an axis use the Rodrigues formula of our own
// (the reference uses Eigen's AngleAxis); the scene's values are not
// bit-identical to the reference's, only its construction is the same.
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#include <algorithm>
#include <limits>
#include <vector>

namespace orc_scene {

struct V3 {
  float x, y, z;
};
struct M3 {
  float m[3][3];
};
struct Iso {
  M3 R;
  V3 t;
};

static V3 v3(float x, float y, float z) { return V3{x, y, z}; }
static V3 add(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static V3 sub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static V3 scale(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
static float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static V3 mul(const M3& R, V3 p) {
  return v3((R.m[0][0] * p.x + R.m[0][1] * p.y) + R.m[0][2] * p.z,
            (R.m[1][0] * p.x + R.m[1][1] * p.y) + R.m[1][2] * p.z,
            (R.m[2][0] * p.x + R.m[2][1] * p.y) + R.m[2][2] * p.z);
}
static V3 mul_t(const M3& R, V3 p) {
  return v3((R.m[0][0] * p.x + R.m[1][0] * p.y) + R.m[2][0] * p.z,
            (R.m[0][1] * p.x + R.m[1][1] * p.y) + R.m[2][1] * p.z,
            (R.m[0][2] * p.x + R.m[1][2] * p.y) + R.m[2][2] * p.z);
}
static M3 mulm(const M3& A, const M3& B) {
  M3 C;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) C.m[r][c] = (A.m[r][0] * B.m[0][c] + A.m[r][1] * B.m[1][c]) + A.m[r][2] * B.m[2][c];
  return C;
}
static Iso compose(const Iso& A, const Iso& B) { return Iso{mulm(A.R, B.R), add(mul(A.R, B.t), A.t)}; }
static V3 apply(const Iso& X, V3 p) { return add(mul(X.R, p), X.t); }
static M3 eye() {
  M3 I{};
  I.m[0][0] = I.m[1][1] = I.m[2][2] = 1.0f;
  return I;
}

// Rotation by `angle` about the unit axis `n`: I + sin(a) S + (1 - cos(a)) S^2.
static M3 angle_axis(float angle, V3 n) {
  const float s = sinf(angle), c = cosf(angle);
  const float S[3][3] = {{0.0f, -n.z, n.y}, {n.z, 0.0f, -n.x}, {-n.y, n.x, 0.0f}};
  M3 R;
  for (int r = 0; r < 3; ++r)
    for (int k = 0; k < 3; ++k) {
      const float s2 = (S[r][0] * S[0][k] + S[r][1] * S[1][k]) + S[r][2] * S[2][k];
      R.m[r][k] = ((r == k ? 1.0f : 0.0f) + s * S[r][k]) + (1.0f - c) * s2;
    }
  return R;
}

// GeometryTraits_<3>::expmap (geometry_3d_impl_.h:179-184): translation is the
// head, rotation rodrigues(tail) (identity below 1e-9 rad).
static Iso expmap(const float* v) {
  Iso X;
  X.t = v3(v[0], v[1], v[2]);
  const V3 w = v3(v[3], v[4], v[5]);
  const float a = sqrtf(dot(w, w));
  X.R = a < 1e-9f ? eye() : angle_axis(a, scale(w, 1.0f / a));
  return X;
}

// Eigen 3.4 random_default_impl<float>::run() on [-1, 1].
static float eigen_random() {
  const float x = -1.0f, y = 1.0f;
  return x + (y - x) * (float)rand() / (float)RAND_MAX;
}

static void reset_generators() {
  srand(1);
  unsigned short zero[3] = {0, 0, 0};
  seed48(zero);
}

// addItem (typedefs.cpp:9-21, 37-49): a square of side
// 2 * width on the local z = 0 plane, (width^2 / resolution^2) points.
static void add_item(std::vector<V3>& dest, float resolution, float width, const Iso& origin) {
  const size_t num_points = (size_t)((width * width) / (resolution * resolution));
  for (size_t k = 0; k < num_points; ++k) {
    // VectorType::Random() * width: coefficients 0, 1, 2 in order; z dropped
    const float px = eigen_random() * width;
    const float py = eigen_random() * width;
    (void)eigen_random();
    dest.push_back(apply(origin, v3(px, py, 0.0f)));
  }
}

// generateScene (typedefs.cpp:23-34, 51-64).
static void generate_scene(std::vector<V3>& dest, float range, float resolution, int num_planes) {
  for (int i = 0; i < num_planes; ++i) {
    const float plane_side = (float)(range * 0.3);
    float pert[6];
    for (int k = 0; k < 6; ++k) pert[k] = eigen_random();  // PerturbationPoseType::Random()
    for (int k = 0; k < 3; ++k) pert[k] *= range;
    const float nrm = sqrtf((pert[3] * pert[3] + pert[4] * pert[4]) + pert[5] * pert[5]);
    for (int k = 3; k < 6; ++k) pert[k] /= nrm;
    const float d = (float)drand48();
    for (int k = 3; k < 6; ++k) pert[k] *= d;
    add_item(dest, resolution, plane_side, expmap(pert));
  }
}

struct ScannerParams {  // scanner.h:6-16
  float min_range = 0.1f;
  float max_range = 200.f;
  float alpha_max = (float)M_PI;
  float alpha_min = (float)-M_PI;
  float phi_max = (float)(M_PI / 4);
  float phi_min = (float)(-M_PI / 4);
  float scan_time = 0.1f;
  int num_h_samples = 1024;
  int num_v_samples = 128;
};

// Scanner::makeScan (scanner.cpp:5-83).  Points are returned in the sensor
// frame at their column's time (azimuth rotation only), each with the
// column's stamp: world = start_pose * [R(w t) | v t] * point.
static void make_scan(std::vector<V3>& out, std::vector<double>& stamps, const std::vector<V3>& scene,
                      const Iso& start_pose, const float* velocities, double t_start, const ScannerParams& P) {
  const float h_fov = P.alpha_max - P.alpha_min;
  const float scan_h_aperture = h_fov / P.num_h_samples;
  const float v_fov = P.phi_max - P.phi_min;
  const float scan_v_aperture = v_fov / P.num_v_samples;
  const float scan_v_ires = 1.f / scan_v_aperture;
  const float beam_time = P.scan_time / P.num_h_samples;
  const float d2_max = P.max_range * P.max_range;
  const float d2_min = P.min_range * P.min_range;
  const V3 uz = v3(0.0f, 0.0f, 1.0f);
  const M3 pi_left = angle_axis(0.5f * scan_h_aperture, uz);
  const M3 pi_right = angle_axis(-0.5f * scan_h_aperture, uz);
  std::vector<V3> v_rays((size_t)P.num_v_samples);
  for (int i = 0; i < P.num_v_samples; ++i)
    v_rays[i] = mul(angle_axis(i * scan_v_aperture + P.phi_min, v3(0.0f, -1.0f, 0.0f)), v3(1.0f, 0.0f, 0.0f));
  const V3 tv = v3(velocities[0], velocities[1], velocities[2]);
  const V3 rv = v3(velocities[3], velocities[4], velocities[5]);
  float r_norm = sqrtf(dot(rv, rv));
  V3 r_axis = uz;
  if (r_norm > 1e-6f)
    r_axis = scale(rv, 1.0f / r_norm);
  else
    r_norm = 0;
  std::vector<float> v_buffer((size_t)P.num_v_samples);
  for (int i = 0; i < P.num_h_samples; ++i) {
    const double integrated_time = (double)beam_time * i;
    const double stamp = t_start + integrated_time;
    const Iso sensor_origin{angle_axis(P.alpha_min + scan_h_aperture * i, uz), v3(0.0f, 0.0f, 0.0f)};
    const Iso sensor_motion{angle_axis((float)(r_norm * integrated_time), r_axis), scale(tv, (float)integrated_time)};
    const Iso world_origin = compose(compose(start_pose, sensor_motion), sensor_origin);
    const V3 n_left = mul(world_origin.R, v3(pi_left.m[0][1], pi_left.m[1][1], pi_left.m[2][1]));
    const V3 n_right = mul(world_origin.R, v3(pi_right.m[0][1], pi_right.m[1][1], pi_right.m[2][1]));
    std::fill(v_buffer.begin(), v_buffer.end(), std::numeric_limits<float>::max());
    for (const V3& s : scene) {
      const V3 dp = sub(s, world_origin.t);
      const float d2 = dot(dp, dp);
      if (d2 < d2_min || d2 > d2_max) continue;
      if (dot(n_left, dp) > 0) continue;
      if (dot(n_right, dp) < 0) continue;
      const V3 ps = mul_t(world_origin.R, dp);
      const float elevation = atan2f(ps.z, ps.x);
      if (elevation < P.phi_min || elevation > P.phi_max) continue;
      const int bin = (int)((elevation - P.phi_min) * scan_v_ires + 0.5f);
      if (bin < 0 || bin >= P.num_v_samples) continue;
      v_buffer[bin] = std::min(v_buffer[bin], d2);
    }
    for (int j = 0; j < P.num_v_samples; ++j) {
      const float range2 = v_buffer[j];
      if (range2 < d2_min || range2 > d2_max) continue;
      out.push_back(mul(sensor_origin.R, scale(v_rays[j], sqrtf(range2))));
      stamps.push_back(stamp);
    }
  }
}

}  // namespace orc_scene

extern "C" {

// generateScene(range, resolution, num_planes) from a fresh process's random
// state.  Writes up to `cap` points (3 floats each) and returns the scene's
// size (which may exceed cap: call with cap = 0 to size the buffer).
int64_t orc_generate_scene(float range, float resolution, int num_planes, float* xyz, int64_t cap) {
  using namespace orc_scene;
  reset_generators();
  std::vector<V3> pts;
  generate_scene(pts, range, resolution, num_planes);
  const int64_t n = (int64_t)pts.size();
  for (int64_t i = 0; i < n && i < cap; ++i) {
    xyz[3 * i + 0] = pts[i].x;
    xyz[3 * i + 1] = pts[i].y;
    xyz[3 * i + 2] = pts[i].z;
  }
  return n;
}

// Scanner::makeScan with the default ScannerParams over `scene` (n points),
// from start_pose (3x4 row-major) under constant velocities [t; w].  Writes up
// to `cap` points and stamps and returns the scan's size.
int64_t orc_make_scan(const float* scene_xyz, int64_t n, const float* start_pose34, const float* velocities,
                      double t_start, float* xyz, double* stamps, int64_t cap) {
  using namespace orc_scene;
  std::vector<V3> scene((size_t)n);
  for (int64_t i = 0; i < n; ++i) scene[i] = v3(scene_xyz[3 * i], scene_xyz[3 * i + 1], scene_xyz[3 * i + 2]);
  Iso X;
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) X.R.m[r][c] = start_pose34[r * 4 + c];
  }
  X.t = v3(start_pose34[3], start_pose34[7], start_pose34[11]);
  std::vector<V3> out;
  std::vector<double> st;
  make_scan(out, st, scene, X, velocities, t_start, ScannerParams());
  const int64_t m = (int64_t)out.size();
  for (int64_t i = 0; i < m && i < cap; ++i) {
    xyz[3 * i + 0] = out[i].x;
    xyz[3 * i + 1] = out[i].y;
    xyz[3 * i + 2] = out[i].z;
    stamps[i] = st[i];
  }
  return m;
}

// GeometryTraits_<3>::expmap in FP32 (3x4 row-major out).
void orc_scene_expmap(const float* v, float* X34) {
  using namespace orc_scene;
  const Iso X = expmap(v);
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) X34[r * 4 + c] = X.R.m[r][c];
  }
  X34[3] = X.t.x;
  X34[7] = X.t.y;
  X34[11] = X.t.z;
}

}  // extern "C"
