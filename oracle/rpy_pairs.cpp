// rpy_pairs.cpp -- oracle (test infrastructure only): O(N^2) FP64 RPY sums.
//
// Restates the reference pair formulas, one target particle at a time, in
// the reference's own term order (tt, then rr, then tr) so that rounding is
// close to the reference's.  The reference kernels compute, for every
// (target i, source j) pair including i == j:
//   r = x_i - x_j, r_hat = r / (|r| + 1e-20)         cuda_functions.cu:302-309
// and branch on |r| against the particle radii.
#include "oracle.h"
#include <cmath>
#include <algorithm>
#include <omp.h>

namespace {

const double kPi = 3.14159265358979323846264338327950288;

struct Geo {
  double rx, ry, rz;  // unit vector (zero at r == 0)
  double r, rm1;
};

inline Geo geo(const double *xi, const double *xj) {
  Geo g;
  g.rx = xi[0] - xj[0];
  g.ry = xi[1] - xj[1];
  g.rz = xi[2] - xj[2];
  g.r = std::sqrt(g.rx * g.rx + g.ry * g.ry + g.rz * g.rz);
  g.rm1 = 1.0 / (g.r + 1e-20);
  g.rx *= g.rm1; g.ry *= g.rm1; g.rz *= g.rm1;
  return g;
}

// Equal-radius translation-translation scalars (A, B), with
// v += A f + B (f . r_hat) r_hat.   cuda_functions.cu:317-350
inline void tt_equal(const Geo &g, double a, double mu, double &A, double &B) {
  if (g.r > 2.0 * a) {
    const double s = 2.0 * a * a * g.rm1 * g.rm1;
    const double c = g.rm1 / (8.0 * kPi * mu);
    A = (1.0 + s / 3.0) * c;
    B = (1.0 - s) * c;
  } else {
    const double b = 3.0 * g.r / (32.0 * a);
    const double c = 1.0 / (6.0 * kPi * mu * a);
    A = (1.0 - 3.0 * b) * c;
    B = b * c;
  }
}

// Equal-radius rotation-rotation: w += A tau + B (tau . r_hat) r_hat.
// cuda_functions.cu:356-386
inline void rr_equal(const Geo &g, double a, double mu, double &A, double &B) {
  if (g.r > 2.0 * a) {
    const double c = -g.rm1 * g.rm1 * g.rm1 / (16.0 * kPi * mu);
    A = c;
    B = -3.0 * c;
  } else {
    const double t = g.r / a;
    const double t3 = t * t * t;
    const double c = 1.0 / (8.0 * kPi * mu * a * a * a);
    A = (1.0 - 27.0 * t / 32.0 + 5.0 * t3 / 64.0) * c;
    B = (9.0 * t / 32.0 - 3.0 * t3 / 64.0) * c;
  }
}

// Equal-radius translation-rotation scalar g:
// w_i += g (f_j x r_hat), v_i += g (tau_j x r_hat).  cuda_functions.cu:392-408
inline double tr_equal(const Geo &g, double a, double mu) {
  if (g.r > 2.0 * a) return g.rm1 * g.rm1 / (8.0 * kPi * mu);
  return g.r * (1.0 - 3.0 * g.r / (8.0 * a)) / (16.0 * kPi * mu * a * a * a);
}

// Unequal-radius tt (seg radius as, blob radius ab).  cuda_functions.cu:608-646
inline void tt_unequal(const Geo &g, double as, double ab, double mu,
                       double &A, double &B) {
  const double dlt = std::fabs(as - ab);
  if (g.r > as + ab) {
    const double s = g.rm1 * g.rm1 * (as * as + ab * ab);
    const double c = g.rm1 / (8.0 * kPi * mu);
    A = (1.0 + s / 3.0) * c;
    B = (1.0 - s) * c;
  } else if (g.r > dlt) {
    const double r3x32 = 32.0 * g.r * g.r * g.r;
    const double p = dlt * dlt + 3.0 * g.r * g.r;
    const double q = dlt * dlt - g.r * g.r;
    const double c = 1.0 / (6.0 * kPi * mu * as * ab);
    A = (0.5 * (as + ab) - p * p / r3x32) * c;
    B = (3.0 * q * q / r3x32) * c;
  } else {
    A = 1.0 / (6.0 * kPi * mu * std::max(as, ab));
    B = 0.0;
  }
}

// Unequal-radius tr scalar (first radius = segment, the rotating particle).
// cuda_functions.cu:652-667
inline double tr_unequal(const Geo &g, double as, double ab, double mu) {
  const double dlt = std::fabs(as - ab);
  if (g.r > as + ab) return g.rm1 * g.rm1 / (8.0 * kPi * mu);
  if (g.r > dlt) {
    const double s = as - ab + g.r;
    return s * s * g.rm1 * g.rm1 *
           (ab * ab + 2.0 * ab * (g.r + as) - 3.0 * (as - g.r) * (as - g.r)) /
           (128.0 * kPi * mu * ab * as * as * as);
  }
  return (as > ab ? 1.0 : 0.0) * g.r / (8.0 * kPi * mu * as * as * as);
}

}  // namespace

extern "C" void orc_Mss(double *V, const double *F, const double *X, int ns,
                        int i0, int i1, double a, double mu, int nthreads) {
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
  for (int i = i0; i < i1; i++) {
    double v[3] = {0, 0, 0}, w[3] = {0, 0, 0};
    for (int j = 0; j < ns; j++) {
      const Geo g = geo(&X[3 * i], &X[3 * j]);
      const double *f = &F[6 * j], *t = &F[6 * j + 3];
      double A, B;
      tt_equal(g, a, mu, A, B);
      double d = g.rx * f[0] + g.ry * f[1] + g.rz * f[2];
      v[0] += A * f[0] + B * d * g.rx;
      v[1] += A * f[1] + B * d * g.ry;
      v[2] += A * f[2] + B * d * g.rz;
      rr_equal(g, a, mu, A, B);
      d = g.rx * t[0] + g.ry * t[1] + g.rz * t[2];
      w[0] += A * t[0] + B * d * g.rx;
      w[1] += A * t[1] + B * d * g.ry;
      w[2] += A * t[2] + B * d * g.rz;
      const double c = tr_equal(g, a, mu);
      w[0] += c * (f[1] * g.rz - f[2] * g.ry);
      w[1] += c * (f[2] * g.rx - f[0] * g.rz);
      w[2] += c * (f[0] * g.ry - f[1] * g.rx);
      v[0] += c * (t[1] * g.rz - t[2] * g.ry);
      v[1] += c * (t[2] * g.rx - t[0] * g.rz);
      v[2] += c * (t[0] * g.ry - t[1] * g.rx);
    }
    for (int k = 0; k < 3; k++) { V[6 * i + k] = v[k]; V[6 * i + 3 + k] = w[k]; }
  }
}

extern "C" void orc_Mbb(double *V, const double *F, const double *X, int nb,
                        int i0, int i1, double a, double mu, int nthreads) {
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
  for (int i = i0; i < i1; i++) {
    double v[3] = {0, 0, 0};
    for (int j = 0; j < nb; j++) {
      const Geo g = geo(&X[3 * i], &X[3 * j]);
      const double *f = &F[3 * j];
      double A, B;
      tt_equal(g, a, mu, A, B);
      const double d = g.rx * f[0] + g.ry * f[1] + g.rz * f[2];
      v[0] += A * f[0] + B * d * g.rx;
      v[1] += A * f[1] + B * d * g.ry;
      v[2] += A * f[2] + B * d * g.rz;
    }
    for (int k = 0; k < 3; k++) V[3 * i + k] = v[k];
  }
}

extern "C" void orc_Msb(double *V, const double *Fb, const double *Xs, int ns,
                        const double *Xb, int nb, int i0, int i1, double as,
                        double ab, double mu, int nthreads) {
  (void)ns;  // targets are [i0, i1)
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
  for (int i = i0; i < i1; i++) {
    double v[3] = {0, 0, 0}, w[3] = {0, 0, 0};
    for (int j = 0; j < nb; j++) {
      const Geo g = geo(&Xs[3 * i], &Xb[3 * j]);
      const double *f = &Fb[3 * j];
      double A, B;
      tt_unequal(g, as, ab, mu, A, B);
      const double d = g.rx * f[0] + g.ry * f[1] + g.rz * f[2];
      v[0] += A * f[0] + B * d * g.rx;
      v[1] += A * f[1] + B * d * g.ry;
      v[2] += A * f[2] + B * d * g.rz;
      const double c = tr_unequal(g, as, ab, mu);
      w[0] += c * (f[1] * g.rz - f[2] * g.ry);
      w[1] += c * (f[2] * g.rx - f[0] * g.rz);
      w[2] += c * (f[0] * g.ry - f[1] * g.rx);
    }
    for (int k = 0; k < 3; k++) { V[6 * i + k] += v[k]; V[6 * i + 3 + k] += w[k]; }
  }
}

extern "C" void orc_Mbs(double *V, const double *Fs, const double *Xb, int nb,
                        const double *Xs, int ns, int i0, int i1, double as,
                        double ab, double mu, int nthreads) {
  (void)nb;  // targets are [i0, i1)
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
  for (int i = i0; i < i1; i++) {
    double v[3] = {0, 0, 0};
    for (int j = 0; j < ns; j++) {
      const Geo g = geo(&Xb[3 * i], &Xs[3 * j]);
      const double *f = &Fs[6 * j], *t = &Fs[6 * j + 3];
      double A, B;
      tt_unequal(g, as, ab, mu, A, B);
      const double d = g.rx * f[0] + g.ry * f[1] + g.rz * f[2];
      v[0] += A * f[0] + B * d * g.rx;
      v[1] += A * f[1] + B * d * g.ry;
      v[2] += A * f[2] + B * d * g.rz;
      const double c = tr_unequal(g, as, ab, mu);
      v[0] += c * (t[1] * g.rz - t[2] * g.ry);
      v[1] += c * (t[2] * g.rx - t[0] * g.rz);
      v[2] += c * (t[0] * g.ry - t[1] * g.rx);
    }
    for (int k = 0; k < 3; k++) V[3 * i + k] = v[k];
  }
}

// cuda_functions.cu:961-1157.  Sources are all segments then all blobs.
extern "C" void orc_barrier(double *f_segs, double *f_blobs_rep,
                            const double *xs, int ns, const double *xb, int nb,
                            int nseg, double as, double ab, double K,
                            int nthreads) {
  auto src = [&](int j) -> const double * {
    return j < ns ? &xs[3 * j] : &xb[3 * (j - ns)];
  };
  const int ntot = ns + nb;
#pragma omp parallel for schedule(static) num_threads(nthreads)
  for (int i = 0; i < ns; i++) {
    double F[3] = {0, 0, 0};
    const int fili = i / nseg;
    for (int j = 0; j < ntot; j++) {
      const double asum = (j < ns) ? 2.0 * as : as + ab;
      const double *xj = src(j);
      const double dx = xs[3 * i] - xj[0], dy = xs[3 * i + 1] - xj[1],
                   dz = xs[3 * i + 2] - xj[2];
      const double dist = std::sqrt(dx * dx + dy * dy + dz * dz);
      const bool excluded = (fili == j / nseg) && std::abs(i - j) <= 1;
      if (!excluded && dist < 1.1 * asum) {
        double fac = std::min(1.0, 1.0 - (10.0 / asum) * (dist - asum));
        fac = K * fac * fac * fac * fac;
        F[0] += fac * dx / dist; F[1] += fac * dy / dist; F[2] += fac * dz / dist;
      }
    }
    for (int k = 0; k < 3; k++) f_segs[6 * i + k] += F[k];
  }
#pragma omp parallel for schedule(static) num_threads(nthreads)
  for (int m = 0; m < nb; m++) {
    double F[3] = {0, 0, 0};
    for (int j = 0; j < ntot; j++) {
      if (j == m + ns) continue;
      const double asum = (j < ns) ? as + ab : 2.0 * ab;
      const double *xj = src(j);
      const double dx = xb[3 * m] - xj[0], dy = xb[3 * m + 1] - xj[1],
                   dz = xb[3 * m + 2] - xj[2];
      const double dist = std::sqrt(dx * dx + dy * dy + dz * dz);
      if (dist < 1.1 * asum) {
        double fac = std::min(1.0, 1.0 - (10.0 / asum) * (dist - asum));
        fac = K * fac * fac * fac * fac;
        F[0] += fac * dx / dist; F[1] += fac * dy / dist; F[2] += fac * dz / dist;
      }
    }
    for (int k = 0; k < 3; k++) f_blobs_rep[3 * m + k] = F[k];
  }
}

// Dense (v_i, w_i) <- (f_j, tau_j) block of the equal-radius mobility, built
// by applying the pair formulas to unit force/torque vectors.
extern "C" void orc_ss_block(double *out, const double *xi, const double *xj,
                             double a, double mu) {
  const Geo g = geo(xi, xj);
  const double rh[3] = {g.rx, g.ry, g.rz};
  double At, Bt, Ar, Br;
  tt_equal(g, a, mu, At, Bt);
  rr_equal(g, a, mu, Ar, Br);
  const double c = tr_equal(g, a, mu);
  for (int k = 0; k < 36; k++) out[k] = 0.0;
  // [r_hat]x such that (e x r_hat) = -[r_hat]x e ; row p of (e_q x r_hat)
  auto cross_e = [&](int p, int q) {
    // (e_q x r)_p = eps_{p q k} r_k
    const int k = 3 - p - q;
    if (p == q) return 0.0;
    const double s = ((p + 1) % 3 == q) ? 1.0 : -1.0;
    return s * rh[k];
  };
  for (int p = 0; p < 3; p++)
    for (int q = 0; q < 3; q++) {
      const double id = (p == q) ? 1.0 : 0.0;
      out[6 * p + q] = At * id + Bt * rh[p] * rh[q];              // v <- f
      out[6 * (p + 3) + (q + 3)] = Ar * id + Br * rh[p] * rh[q];  // w <- tau
      out[6 * (p + 3) + q] = c * cross_e(p, q);                   // w <- f
      out[6 * p + (q + 3)] = c * cross_e(p, q);                   // v <- tau
    }
}
