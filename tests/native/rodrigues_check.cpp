// Checks kdh_solve.h's FP64 rotation (rodrigues_d) on the host: against the
// matrix exponential of [w]x summed as a power series in long double (an
// independent algorithm), for angles across the small-angle (Taylor) branch,
// its boundary at 0.25 rad and the closed form beyond; plus R^T R = I,
// det R = 1 and R w = w.  Exit status 0 when every check holds.
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "../../kd_slam2_amd/csrc/kdh_solve.h"

using namespace kdh;

static void expm_series(const double w[3], long double R[9]) {
  const long double S[9] = {0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0};
  long double term[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  for (int i = 0; i < 9; ++i) R[i] = term[i];
  for (int k = 1; k < 60; ++k) {
    long double next[9];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        long double s = 0;
        for (int m = 0; m < 3; ++m) s += term[r * 3 + m] * S[m * 3 + c];
        next[r * 3 + c] = s / k;
      }
    for (int i = 0; i < 9; ++i) {
      term[i] = next[i];
      R[i] += term[i];
    }
  }
}

int main() {
  srand(7);
  double worst_series = 0, worst_orth = 0, worst_det = 0, worst_axis = 0;
  int n_small = 0, n_large = 0;
  const double scales[] = {0.0, 1e-12, 1e-6, 1e-3, 0.05, 0.2, 0.2499, 0.25, 0.2501, 0.6, 1.5, 3.0};
  for (double sc : scales) {
    for (int k = 0; k < 2000; ++k) {
      double w[3];
      for (int i = 0; i < 3; ++i) w[i] = rand() / (double)RAND_MAX - 0.5;
      const double nw = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
      for (int i = 0; i < 3; ++i) w[i] *= sc / nw;
      (w[0] * w[0] + w[1] * w[1] + w[2] * w[2] < 0.0625 ? n_small : n_large)++;
      const M3d R = rodrigues_d(v3(w[0], w[1], w[2]));
      long double E[9];
      expm_series(w, E);
      for (int i = 0; i < 9; ++i) worst_series = std::fmax(worst_series, std::fabs((double)(R.m[i] - E[i])));
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
          double s = 0;
          for (int m = 0; m < 3; ++m) s += R.m[m * 3 + r] * R.m[m * 3 + c];
          worst_orth = std::fmax(worst_orth, std::fabs(s - (r == c ? 1.0 : 0.0)));
        }
      const double det = R.m[0] * (R.m[4] * R.m[8] - R.m[5] * R.m[7]) - R.m[1] * (R.m[3] * R.m[8] - R.m[5] * R.m[6]) +
                         R.m[2] * (R.m[3] * R.m[7] - R.m[4] * R.m[6]);
      worst_det = std::fmax(worst_det, std::fabs(det - 1.0));
      for (int r = 0; r < 3; ++r) {
        const double s = (R.m[r * 3 + 0] * w[0] + R.m[r * 3 + 1] * w[1]) + R.m[r * 3 + 2] * w[2];
        worst_axis = std::fmax(worst_axis, std::fabs(s - w[r]));
      }
    }
  }
  std::printf("small %d large %d: |R - expm| %.3g, |R^T R - I| %.3g, |det - 1| %.3g, |R w - w| %.3g\n", n_small,
              n_large, worst_series, worst_orth, worst_det, worst_axis);
  const bool ok = worst_series < 2e-15 && worst_orth < 4e-15 && worst_det < 4e-15 && worst_axis < 2e-15;
  return ok && n_small > 0 && n_large > 0 ? 0 : 1;
}
