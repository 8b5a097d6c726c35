// seed.cpp -- CPU restatement of the reference's unit-sphere seeding
// (seeding.cu:13-389, cuda_functions.cu:1159-1250).  TEST INFRASTRUCTURE ONLY.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace {

// seeding.cu:19-38 / :271-290
void spiral(int N, double *X) {
  double phi = 0.0;
  for (int n = 0; n < N; n++) {
    X[3 * n + 2] = (N == 1) ? -1.0 : 2.0 * n / double(N - 1) - 1.0;
    const double r = std::sqrt(1.0 - X[3 * n + 2] * X[3 * n + 2]);
    if (n == 0 || n == N - 1) {
      phi = 0.0;
    } else {
      phi += 3.6 / (r * std::sqrt(double(N)));
    }
    X[3 * n] = r * std::cos(phi);
    X[3 * n + 1] = r * std::sin(phi);
  }
}

}  // namespace

extern "C" {

// Uniform seeding.  order 1: forces summed per particle over j in index
// order (the reference's GPU kernel, cuda_functions.cu:1159-1250, used for
// N >= 1000); order 0: the reference's CPU loop over pairs n < m, adding to
// both (seeding.cu:66-90, used for N < 1000).  Returns the iterations, or
// -1 if max_it was reached first.
int orc_seed_uniform(int N, int order, int max_it, double *X) {
  spiral(N, X);
  std::vector<double> F(3 * size_t(N));
  double biggest = 1e4;
  int it = 0;
  while (biggest > 1e-4) {
    if (it == max_it) return -1;
    biggest = 0.0;
    std::fill(F.begin(), F.end(), 0.0);
    if (order == 0) {
      for (int n = 0; n < N - 1; n++) {
        const double Xn[3] = {X[3 * n], X[3 * n + 1], X[3 * n + 2]};
        for (int m = n + 1; m < N; m++) {
          const double d[3] = {Xn[0] - X[3 * m], Xn[1] - X[3 * m + 1], Xn[2] - X[3 * m + 2]};
          const double dn = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
          for (int k = 0; k < 3; k++) {
            F[3 * n + k] += 5.0 * (2.0 - dn) * d[k] / dn;
            F[3 * m + k] -= 5.0 * (2.0 - dn) * d[k] / dn;
          }
        }
      }
    } else {
      for (int i = 0; i < N; i++) {
        double f[3] = {0, 0, 0};
        for (int j = 0; j < N; j++) {
          if (j == i) continue;
          const double d[3] = {X[3 * i] - X[3 * j], X[3 * i + 1] - X[3 * j + 1], X[3 * i + 2] - X[3 * j + 2]};
          const double dist = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
          const double dm1 = 1.0 / dist;
          const double fac = 5.0 * (2.0 - dist);
          for (int k = 0; k < 3; k++) f[k] += fac * d[k] * dm1;
        }
        for (int k = 0; k < 3; k++) F[3 * i + k] = f[k];
      }
    }
    for (int n = 0; n < N; n++) {
      double p[3] = {X[3 * n] + F[3 * n], X[3 * n + 1] + F[3 * n + 1], X[3 * n + 2] + F[3 * n + 2]};
      const double nrm = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
      for (int k = 0; k < 3; k++) {
        p[k] /= nrm;
        biggest = std::max(std::abs(p[k] - X[3 * n + k]), biggest);
        X[3 * n + k] = p[k];
      }
    }
    it++;
  }
  return it;
}

// Equatorial (seeding.cu:140-175) and platy (seeding.cu:177-230) bands.
void orc_seed_band(int N, int platy, double r_over_l, double *X) {
  const double theta_max = 0.5 / r_over_l;
  const double h = 2.0 * std::sin(theta_max);
  const int in_band = platy ? int(std::round(0.9 * N)) : N;
  const int total = int(std::ceil(in_band * 2.0 / h));
  const int offset = int(std::round(0.5 * (total - in_band)));
  double phi = 0.0;
  for (int n = 0; n < in_band; n++) {
    X[3 * n + 2] = 2.0 * (n + offset) / double(total - 1) - 1.0;
    const double r = std::sqrt(1.0 - X[3 * n + 2] * X[3 * n + 2]);
    X[3 * n] = r * std::cos(phi);
    X[3 * n + 1] = r * std::sin(phi);
    phi += 3.6 / (r * std::sqrt(double(total)));
  }
  for (int n = in_band; n < N; n++) {
    X[3 * n + 2] = -0.9;
    const double r = std::sqrt(1.0 - X[3 * n + 2] * X[3 * n + 2]);
    phi = 2.0 * 3.14159265358979323846 * double(n - in_band) / double(N - in_band);
    X[3 * n] = r * std::cos(phi);
    X[3 * n + 1] = r * std::sin(phi);
  }
}

}  // extern "C"
