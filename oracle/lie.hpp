// lie.hpp -- oracle (test infrastructure only): quaternion / SO(3) helpers
// and small dense matrices, restating quaternion.cpp and matrix.cpp.
#pragma once
#include <cmath>
#include <vector>
#include <cstring>

namespace orc {

constexpr double kPi = 3.14159265358979323846264338327950288;

struct Quat {
  double s = 1, v[3] = {0, 0, 0};
};

// q1 *= q2 semantics (quaternion.cpp:104-121): returns p*q.
inline Quat qmul(const Quat &p, const Quat &q) {
  Quat r;
  r.s = p.s * q.s - (p.v[0] * q.v[0] + p.v[1] * q.v[1] + p.v[2] * q.v[2]);
  r.v[0] = p.s * q.v[0] + q.s * p.v[0] + p.v[1] * q.v[2] - p.v[2] * q.v[1];
  r.v[1] = p.s * q.v[1] + q.s * p.v[1] + p.v[2] * q.v[0] - p.v[0] * q.v[2];
  r.v[2] = p.s * q.v[2] + q.s * p.v[2] + p.v[0] * q.v[1] - p.v[1] * q.v[0];
  return r;
}
inline Quat qconj(Quat q) { q.v[0] = -q.v[0]; q.v[1] = -q.v[1]; q.v[2] = -q.v[2]; return q; }
inline Quat qsub(const Quat &a, const Quat &b) {
  Quat r; r.s = a.s - b.s;
  for (int k = 0; k < 3; k++) r.v[k] = a.v[k] - b.v[k];
  return r;
}
inline Quat qscale(Quat q, double f) { q.s *= f; for (int k = 0; k < 3; k++) q.v[k] *= f; return q; }

// quaternion.cpp:180-210
inline Quat qsqrt(Quat q) {
  if (q.s < -0.999999999) {
    Quat r; r.s = 0; r.v[0] = 0; r.v[1] = 0; r.v[2] = 1;
    return r;
  }
  q.s = std::sqrt(0.5 * (1.0 + q.s));
  const double t = 2.0 * q.s;
  q.v[0] /= t; q.v[1] /= t; q.v[2] /= t;
  return q;
}

// frame vectors = columns of the rotation matrix (quaternion.cpp:218-240)
inline void tangent(const Quat &q, double *t) {
  t[0] = 1.0 - 2.0 * (q.v[1] * q.v[1] + q.v[2] * q.v[2]);
  t[1] = 2.0 * (q.v[0] * q.v[1] + q.s * q.v[2]);
  t[2] = 2.0 * (q.v[0] * q.v[2] - q.s * q.v[1]);
}
inline void normal(const Quat &q, double *n) {
  n[0] = 2.0 * (q.v[0] * q.v[1] - q.s * q.v[2]);
  n[1] = 1.0 - 2.0 * (q.v[0] * q.v[0] + q.v[2] * q.v[2]);
  n[2] = 2.0 * (q.v[1] * q.v[2] + q.s * q.v[0]);
}
inline void binormal(const Quat &q, double *b) {
  b[0] = 2.0 * (q.v[0] * q.v[2] + q.s * q.v[1]);
  b[1] = 2.0 * (q.v[1] * q.v[2] - q.s * q.v[0]);
  b[2] = 1.0 - 2.0 * (q.v[0] * q.v[0] + q.v[1] * q.v[1]);
}

// ---- 3x3 row-major ----
struct M3 {
  double a[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  double &operator()(int i, int j) { return a[3 * i + j]; }
  double operator()(int i, int j) const { return a[3 * i + j]; }
};
inline M3 eye3() { M3 m; m(0, 0) = m(1, 1) = m(2, 2) = 1; return m; }
inline M3 operator*(const M3 &A, const M3 &B) {
  M3 C;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double s = 0;
      for (int k = 0; k < 3; k++) s += A(i, k) * B(k, j);
      C(i, j) = s;
    }
  return C;
}
inline M3 operator+(M3 A, const M3 &B) { for (int k = 0; k < 9; k++) A.a[k] += B.a[k]; return A; }
inline M3 operator-(M3 A, const M3 &B) { for (int k = 0; k < 9; k++) A.a[k] -= B.a[k]; return A; }
inline M3 operator*(double f, M3 A) { for (int k = 0; k < 9; k++) A.a[k] *= f; return A; }
inline M3 transpose(const M3 &A) {
  M3 T;
  for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) T(i, j) = A(j, i);
  return T;
}
inline void mv3(const M3 &A, const double *x, double *y) {
  double t[3];
  for (int i = 0; i < 3; i++) t[i] = A(i, 0) * x[0] + A(i, 1) * x[1] + A(i, 2) * x[2];
  y[0] = t[0]; y[1] = t[1]; y[2] = t[2];
}
// matrix.cpp:665-679: rcross(v) w = w x v
inline M3 rcross(const double *v) {
  M3 M;
  M(0, 1) = v[2]; M(0, 2) = -v[1];
  M(1, 0) = -v[2]; M(1, 2) = v[0];
  M(2, 0) = v[1]; M(2, 1) = -v[0];
  return M;
}
inline void cross(const double *a, const double *b, double *c) {
  const double t0 = a[1] * b[2] - a[2] * b[1], t1 = a[2] * b[0] - a[0] * b[2],
               t2 = a[0] * b[1] - a[1] * b[0];
  c[0] = t0; c[1] = t1; c[2] = t2;
}

// quaternion.cpp:260-302
inline M3 rot_mat(const Quat &q) {
  M3 R;
  R(0, 0) = 1.0 - 2.0 * (q.v[1] * q.v[1] + q.v[2] * q.v[2]);
  R(1, 1) = 1.0 - 2.0 * (q.v[0] * q.v[0] + q.v[2] * q.v[2]);
  R(2, 2) = 1.0 - 2.0 * (q.v[0] * q.v[0] + q.v[1] * q.v[1]);
  R(1, 0) = 2.0 * (q.v[0] * q.v[1] + q.s * q.v[2]);
  R(0, 1) = 2.0 * (q.v[0] * q.v[1] - q.s * q.v[2]);
  R(2, 0) = 2.0 * (q.v[0] * q.v[2] - q.s * q.v[1]);
  R(0, 2) = 2.0 * (q.v[0] * q.v[2] + q.s * q.v[1]);
  R(2, 1) = 2.0 * (q.v[1] * q.v[2] + q.s * q.v[0]);
  R(1, 2) = 2.0 * (q.v[1] * q.v[2] - q.s * q.v[0]);
  return R;
}

// quaternion.cpp:455-478
inline Quat lie_exp(const double *u) {
  const double th = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  Quat q;
  if (th >= 1e-10) {
    const double sn = std::sin(0.5 * th) / th;
    q.s = std::cos(0.5 * th);
    q.v[0] = sn * u[0]; q.v[1] = sn * u[1]; q.v[2] = sn * u[2];
  }
  return q;
}

// quaternion.cpp:488-497: qmid = sqrt(q2 conj(q1)) q1
inline Quat midpoint(const Quat &q1, const Quat &q2) {
  return qmul(qsqrt(qmul(q2, qconj(q1))), q1);
}

constexpr double kSmallAngle = 1e-5;  // quaternion.cpp:7

// quaternion.cpp:507-541
inline void dexp(double *out, const double *u, const double *v) {
  const double th2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  const double th = std::sqrt(th2);
  double alpha, beta;
  if (th < kSmallAngle) {
    const double th4 = th2 * th2;
    beta = (1.0 / 3.0 + th2 / 72.0) / (2.0 + th2 / 6.0 + th4 / 72.0);
    alpha = 1.0 / 6.0 + 2.0 * beta * (1.0 - th2 / 12.0);
  } else {
    alpha = std::sin(0.5 * th);
    alpha *= 2.0 * alpha / th2;
    beta = (th - std::sin(th)) / (th * th2);
  }
  double uv[3], uuv[3];
  cross(u, v, uv);
  cross(u, uv, uuv);
  for (int k = 0; k < 3; k++) out[k] = v[k] + alpha * uv[k] + beta * uuv[k];
}

inline double dexpinv_fac(const double *u) {
  const double th2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  const double th = std::sqrt(th2);
  if (th < kSmallAngle) return -1.0 / 12.0;
  return (0.5 * th * std::cos(0.5 * th) / std::sin(0.5 * th) - 1.0) / th2;
}
// quaternion.cpp:543-566
inline void dexpinv(double *out, const double *u, const double *v) {
  const double fac = dexpinv_fac(u);
  double uv[3], uuv[3];
  cross(u, v, uv);
  cross(u, uv, uuv);
  for (int k = 0; k < 3; k++) out[k] = v[k] - 0.5 * uv[k] - fac * uuv[k];
}
// quaternion.cpp:568-591
inline void dexpinv_transpose(double *out, const double *u, const double *v) {
  const double fac = dexpinv_fac(u);
  double uv[3], uuv[3];
  cross(u, v, uv);
  cross(u, uv, uuv);
  for (int k = 0; k < 3; k++) out[k] = v[k] + 0.5 * uv[k] - fac * uuv[k];
}

// ---- dense row-major matrix ----
struct Mat {
  int r = 0, c = 0;
  std::vector<double> d;
  Mat() {}
  Mat(int r_, int c_) : r(r_), c(c_), d(size_t(r_) * c_, 0.0) {}
  double &operator()(int i, int j) { return d[size_t(i) * c + j]; }
  double operator()(int i, int j) const { return d[size_t(i) * c + j]; }
  void zero() { std::fill(d.begin(), d.end(), 0.0); }
  void set3(int i0, int j0, const M3 &B) {
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) (*this)(i0 + i, j0 + j) = B(i, j);
  }
  void add3(int i0, int j0, const M3 &B) {
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) (*this)(i0 + i, j0 + j) += B(i, j);
  }
  M3 get3(int i0, int j0) const {
    M3 B;
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) B(i, j) = (*this)(i0 + i, j0 + j);
    return B;
  }
};

inline Mat matmul(const Mat &A, const Mat &B) {
  Mat C(A.r, B.c);
  for (int i = 0; i < A.r; i++)
    for (int k = 0; k < A.c; k++) {
      const double a = A(i, k);
      if (a == 0.0) continue;
      for (int j = 0; j < B.c; j++) C(i, j) += a * B(k, j);
    }
  return C;
}

// In-place inverse by Gauss-Jordan elimination with partial pivoting (the
// reference calls LAPACK dgetrf_/dgetri_, matrix.cpp:538-567, and exits on
// a singular matrix).  Returns false if singular.
inline bool invert(Mat &A) {
  const int n = A.r;
  std::vector<int> piv(n);
  for (int i = 0; i < n; i++) piv[i] = i;
  for (int k = 0; k < n; k++) {
    int p = k;
    double best = std::fabs(A(k, k));
    for (int i = k + 1; i < n; i++)
      if (std::fabs(A(i, k)) > best) { best = std::fabs(A(i, k)); p = i; }
    if (best == 0.0) return false;
    if (p != k) {
      for (int j = 0; j < n; j++) std::swap(A(k, j), A(p, j));
      std::swap(piv[k], piv[p]);
    }
    const double inv = 1.0 / A(k, k);
    A(k, k) = 1.0;
    for (int j = 0; j < n; j++) A(k, j) *= inv;
    for (int i = 0; i < n; i++) {
      if (i == k) continue;
      const double f = A(i, k);
      if (f == 0.0) continue;
      A(i, k) = 0.0;
      for (int j = 0; j < n; j++) A(i, j) -= f * A(k, j);
    }
  }
  // undo the column permutation implied by the row swaps
  Mat B(n, n);
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) B(i, piv[j]) = A(i, j);
  A = B;
  return true;
}

}  // namespace orc
