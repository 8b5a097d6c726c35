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

// ---------------------------------------------------------------------------
// Segment-segment mobility above a no-slip plane wall at z = 0 (Swan & Brady
// 2007, as the reference evaluates it in cuda_functions.cu:8-242, including
// its sign convention for the rotation-translation wall terms).  The
// reference assumes RSEG = MU = 1; here positions are divided by a and the
// blocks scaled by 1/(mu a) (tt), 1/(mu a^2) (tr, rt) and 1/(mu a^3) (rr),
// which is the same expression for a = mu = 1.  Pairs j != i use the
// far-field RPY form at any separation, as the reference does.
extern "C" void orc_Mss_wall(double *V, const double *F, const double *X, int ns, int i0, int i1,
                             double a, double mu, int nthreads) {
  const double FF = 1.0 / (6.0 * kPi);   // ForceFactor
  const double TF = 1.0 / (8.0 * kPi);   // TorqueFactor
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
  for (int i = i0; i < i1; i++) {
    // v = vt (from forces) + vr (from torques); w = wt (forces) + wr (torques)
    double vt[3] = {0, 0, 0}, vr[3] = {0, 0, 0}, wt[3] = {0, 0, 0}, wr[3] = {0, 0, 0};
    const double xi = X[3 * i] / a, yi = X[3 * i + 1] / a, zi = X[3 * i + 2] / a;
    for (int j = 0; j < ns; j++) {
      const double *f = &F[6 * j], *t = &F[6 * j + 3];
      const double xj = X[3 * j] / a, yj = X[3 * j + 1] / a, zj = X[3 * j + 2] / a;
      if (j == i) {
        const double hm1 = 1.0 / zi, hm3 = hm1 * hm1 * hm1, hm4 = hm3 * hm1, hm5 = hm4 * hm1;
        const double par = FF - (9.0 * hm1 - 2.0 * hm3 + hm5) * FF / 16.0;
        const double perp = FF - (9.0 * hm1 - 4.0 * hm3 + hm5) * FF / 8.0;
        const double c = 3.0 * FF / 32.0 * hm4;
        vt[0] += par * f[0]; vt[1] += par * f[1]; vt[2] += perp * f[2];
        vr[0] += c * t[1]; vr[1] -= c * t[0];
        const double rpar = TF - 15.0 * FF / 64.0 * hm3;
        const double rperp = TF - 3.0 * FF / 32.0 * hm3;
        wr[0] += rpar * t[0]; wr[1] += rpar * t[1]; wr[2] += rperp * t[2];
        wt[0] -= c * f[1]; wt[1] += c * f[0];
        continue;
      }
      // unbounded RPY, far-field form
      const double dx = xi - xj, dy = yi - yj, dz = zi - zj;
      double rm1 = 1.0 / std::sqrt(dx * dx + dy * dy + dz * dz);
      const double rx = dx * rm1, ry = dy * rm1, rz = dz * rm1;
      const double rf = rx * f[0] + ry * f[1] + rz * f[2];
      const double rt = rx * t[0] + ry * t[1] + rz * t[2];
      double A = TF * rm1, B = 1.0 + 2.0 / 3.0 * rm1 * rm1, C = 1.0 - 2.0 * rm1 * rm1;
      vt[0] += A * (B * f[0] + C * rf * rx);
      vt[1] += A * (B * f[1] + C * rf * ry);
      vt[2] += A * (B * f[2] + C * rf * rz);
      A = 0.5 * TF * rm1 * rm1 * rm1;
      wr[0] += A * (3.0 * rt * rx - t[0]);
      wr[1] += A * (3.0 * rt * ry - t[1]);
      wr[2] += A * (3.0 * rt * rz - t[2]);
      A = TF * rm1 * rm1;
      vr[0] += A * (t[1] * rz - t[2] * ry);
      vr[1] += A * (t[2] * rx - t[0] * rz);
      vr[2] += A * (t[0] * ry - t[1] * rx);
      wt[0] += A * (f[1] * rz - f[2] * ry);
      wt[1] += A * (f[2] * rx - f[0] * rz);
      wt[2] += A * (f[0] * ry - f[1] * rx);
      // wall corrections: image at (xj, yj, -zj)
      const double sz = zi + zj;
      rm1 = 1.0 / std::sqrt(dx * dx + dy * dy + sz * sz);
      const double rm2 = rm1 * rm1, rm3 = rm2 * rm1, rm4 = rm3 * rm1, rm5 = rm4 * rm1;
      double ex = dx * rm1, ey = dy * rm1;
      const double ez = sz * rm1;
      double h = zj / sz;
      const double ef = ex * f[0] + ey * f[1] + ez * f[2];
      const double et = ex * t[0] + ey * t[1] + ez * t[2];
      const double ez2 = ez * ez;
      {  // tt
        const double a0 = -(0.75 * rm1 * (1.0 + 2.0 * h * (1.0 - h) * ez2) +
                            0.5 * (rm3 * (1.0 - 3.0 * ez2) - rm5 * (1.0 - 5.0 * ez2)));
        const double a1 = -(0.75 * rm1 * (1.0 - 6.0 * h * (1.0 - h) * ez2) -
                            1.5 * rm3 * (1.0 - 5.0 * ez2) + 2.5 * rm5 * (1.0 - 7.0 * ez2)) * ef;
        const double a2 = ez * f[2] * (1.5 * h * rm1 * (1.0 - 6.0 * (1.0 - h) * ez2) -
                                       3.0 * rm3 * (1.0 - 5.0 * ez2) + 5.0 * rm5 * (2.0 - 7.0 * ez2));
        const double a3 = (1.5 * h * rm1 - 5.0 * rm5) * ef;
        const double a4 = -((3.0 * h * h * rm1 + 3.0 * rm3) * ez2 + (2.0 - 15.0 * ez2) * rm5) * f[2];
        vt[0] += FF * (a0 * f[0] + (a1 + a2) * ex);
        vt[1] += FF * (a0 * f[1] + (a1 + a2) * ey);
        vt[2] += FF * (a0 * f[2] + (a1 + a2 + a3) * ez + a4);
      }
      {  // rr
        const double a0 = rm3 * (0.375 - 2.25 * ez2);
        const double a1 = -1.125 * rm3 * et;
        const double a2 = 2.25 * ez * rm3 * et;
        const double a3 = 2.25 * rm3 * (ex * t[1] - ey * t[0]);
        wr[0] += FF * (a0 * t[0] + a1 * ex - a3 * ey);
        wr[1] += FF * (a0 * t[1] + a1 * ey + a3 * ex);
        wr[2] += FF * (a0 * t[2] + a1 * ez + a2);
      }
      const double q0 = 0.75 * rm2;
      {  // rt (angular velocity from forces), the reference's sign (:197-204)
        const double a1 = f[2] * (9.0 * h * ez2 * rm2 + (1.5 - 15.0 * ez2) * rm4);
        const double a2 = -ez * ef * (4.5 * h * rm2 - 7.5 * rm4);
        const double a3 = -1.5 * ez * (h * rm2 - rm4);
        wt[0] -= FF * (q0 * (f[1] * ez - f[2] * ey) - (a1 + a2) * ey + a3 * f[1]);
        wt[1] -= FF * (q0 * (f[2] * ex - f[0] * ez) + (a1 + a2) * ex - a3 * f[0]);
        wt[2] -= FF * q0 * (f[0] * ey - f[1] * ex);
      }
      {  // tr: the rt block of the pair (j, i) transposed (:206-224)
        h = 1.0 - h;
        ex = -ex;
        ey = -ey;
        const double ect = ex * t[1] - ey * t[0];
        const double a1 = (9.0 * h * ez2 * rm2 + (1.5 - 15.0 * ez2) * rm4) * ect;
        const double a2 = -ez * (4.5 * h * rm2 - 7.5 * rm4) * ect;
        const double a3 = -1.5 * ez * (h * rm2 - rm4);
        vr[0] -= FF * (q0 * (t[2] * ey - t[1] * ez) + a2 * ex - a3 * t[1]);
        vr[1] -= FF * (q0 * (t[0] * ez - t[2] * ex) + a2 * ey + a3 * t[0]);
        vr[2] -= FF * (q0 * (t[1] * ex - t[0] * ey) + a2 * ez + a1);
      }
    }
    const double s1 = 1.0 / (mu * a), s2 = s1 / a, s3 = s2 / a;
    for (int k = 0; k < 3; k++) {
      V[6 * i + k] = s1 * vt[k] + s2 * vr[k];
      V[6 * i + 3 + k] = s2 * wt[k] + s3 * wr[k];
    }
  }
}
