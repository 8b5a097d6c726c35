// gmi_lie.hpp -- device-side SO(3) / unit-quaternion helpers (FP64).
//
// Quaternions are (s, v0, v1, v2) and rotate body-frame vectors to the lab
// frame.  The functions follow the reference's definitions:
//   product, conjugate, square root        quaternion.cpp:104-121,171-210
//   frame vectors = rotation-matrix columns quaternion.cpp:218-302
//   Lie exponential of u in so(3)          quaternion.cpp:455-478
//   midpoint sqrt(q2 conj(q1)) q1          quaternion.cpp:488-497
//   dexp / dexpinv / dexpinv^T             quaternion.cpp:507-591
//   rcross(v): rcross(v) w = w x v         matrix.cpp:665-679
#pragma once
#include <hip/hip_runtime.h>

namespace gmi {

struct dq { double s, x, y, z; };
struct d3 { double x, y, z; };
struct m3 { double a[9]; };  // row-major

__device__ __forceinline__ d3 mk3(double x, double y, double z) { return {x, y, z}; }
__device__ __forceinline__ d3 operator+(d3 a, d3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ d3 operator-(d3 a, d3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ d3 operator*(double f, d3 a) { return {f * a.x, f * a.y, f * a.z}; }
__device__ __forceinline__ double dot3(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ d3 cross3(d3 a, d3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

__device__ __forceinline__ dq qmul(dq p, dq q) {
  return {p.s * q.s - (p.x * q.x + p.y * q.y + p.z * q.z),
          p.s * q.x + q.s * p.x + p.y * q.z - p.z * q.y,
          p.s * q.y + q.s * p.y + p.z * q.x - p.x * q.z,
          p.s * q.z + q.s * p.z + p.x * q.y - p.y * q.x};
}
__device__ __forceinline__ dq qconj(dq q) { return {q.s, -q.x, -q.y, -q.z}; }
__device__ __forceinline__ dq qsqrt(dq q) {
  if (q.s < -0.999999999) return {0.0, 0.0, 0.0, 1.0};
  const double s = sqrt(0.5 * (1.0 + q.s));
  const double t = 2.0 * s;
  return {s, q.x / t, q.y / t, q.z / t};
}
__device__ __forceinline__ dq qmid(dq q1, dq q2) { return qmul(qsqrt(qmul(q2, qconj(q1))), q1); }

__device__ __forceinline__ d3 qtangent(dq q) {
  return {1.0 - 2.0 * (q.y * q.y + q.z * q.z), 2.0 * (q.x * q.y + q.s * q.z),
          2.0 * (q.x * q.z - q.s * q.y)};
}
__device__ __forceinline__ d3 qnormal(dq q) {
  return {2.0 * (q.x * q.y - q.s * q.z), 1.0 - 2.0 * (q.x * q.x + q.z * q.z),
          2.0 * (q.y * q.z + q.s * q.x)};
}
__device__ __forceinline__ d3 qbinormal(dq q) {
  return {2.0 * (q.x * q.z + q.s * q.y), 2.0 * (q.y * q.z - q.s * q.x),
          1.0 - 2.0 * (q.x * q.x + q.y * q.y)};
}

__device__ __forceinline__ m3 rotm(dq q) {
  const d3 t = qtangent(q), n = qnormal(q), b = qbinormal(q);
  return {{t.x, n.x, b.x, t.y, n.y, b.y, t.z, n.z, b.z}};
}
__device__ __forceinline__ d3 mv(const m3 &A, d3 v) {
  return {A.a[0] * v.x + A.a[1] * v.y + A.a[2] * v.z, A.a[3] * v.x + A.a[4] * v.y + A.a[5] * v.z,
          A.a[6] * v.x + A.a[7] * v.y + A.a[8] * v.z};
}
__device__ __forceinline__ d3 mtv(const m3 &A, d3 v) {  // A^T v
  return {A.a[0] * v.x + A.a[3] * v.y + A.a[6] * v.z, A.a[1] * v.x + A.a[4] * v.y + A.a[7] * v.z,
          A.a[2] * v.x + A.a[5] * v.y + A.a[8] * v.z};
}
__device__ __forceinline__ m3 mm(const m3 &A, const m3 &B) {
  m3 C;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++)
      C.a[3 * i + j] = A.a[3 * i] * B.a[j] + A.a[3 * i + 1] * B.a[3 + j] + A.a[3 * i + 2] * B.a[6 + j];
  return C;
}
__device__ __forceinline__ m3 madd(const m3 &A, const m3 &B) {
  m3 C;
#pragma unroll
  for (int k = 0; k < 9; k++) C.a[k] = A.a[k] + B.a[k];
  return C;
}
__device__ __forceinline__ m3 msub(const m3 &A, const m3 &B) {
  m3 C;
#pragma unroll
  for (int k = 0; k < 9; k++) C.a[k] = A.a[k] - B.a[k];
  return C;
}
__device__ __forceinline__ m3 mscale(double f, const m3 &A) {
  m3 C;
#pragma unroll
  for (int k = 0; k < 9; k++) C.a[k] = f * A.a[k];
  return C;
}
__device__ __forceinline__ m3 mtrans(const m3 &A) {
  return {{A.a[0], A.a[3], A.a[6], A.a[1], A.a[4], A.a[7], A.a[2], A.a[5], A.a[8]}};
}
__device__ __forceinline__ m3 meye() { return {{1, 0, 0, 0, 1, 0, 0, 0, 1}}; }
__device__ __forceinline__ m3 mzero() { return {{0, 0, 0, 0, 0, 0, 0, 0, 0}}; }
__device__ __forceinline__ m3 rcross(d3 v) { return {{0, v.z, -v.y, -v.z, 0, v.x, v.y, -v.x, 0}}; }

__device__ __forceinline__ dq lie_exp(d3 u) {
  const double th = sqrt(dot3(u, u));
  if (th < 1e-10) return {1.0, 0.0, 0.0, 0.0};
  double sn, cs;
  sincos(0.5 * th, &sn, &cs);
  sn /= th;
  return {cs, sn * u.x, sn * u.y, sn * u.z};
}

constexpr double kSmallAngle = 1e-5;

__device__ __forceinline__ double dexpinv_fac(d3 u) {
  const double th2 = dot3(u, u), th = sqrt(th2);
  if (th < kSmallAngle) return -1.0 / 12.0;
  double sn, cs;
  sincos(0.5 * th, &sn, &cs);
  return (0.5 * th * cs / sn - 1.0) / th2;
}
__device__ __forceinline__ d3 dexpinv(d3 u, d3 v) {
  const double f = dexpinv_fac(u);
  const d3 uv = cross3(u, v), uuv = cross3(u, uv);
  return v - 0.5 * uv - f * uuv;
}
__device__ __forceinline__ d3 dexpinv_T(d3 u, d3 v) {
  const double f = dexpinv_fac(u);
  const d3 uv = cross3(u, v), uuv = cross3(u, uv);
  return v + 0.5 * uv - f * uuv;
}

// SoA loads/stores of 3- and 4-vectors (component c at p[c*ld + i]).
__device__ __forceinline__ d3 ld3(const double *p, int ld, int i) {
  return {p[i], p[ld + i], p[2 * ld + i]};
}
__device__ __forceinline__ void st3(double *p, int ld, int i, d3 v) {
  p[i] = v.x; p[ld + i] = v.y; p[2 * ld + i] = v.z;
}
__device__ __forceinline__ dq ld4(const double *p, int ld, int i) {
  return {p[i], p[ld + i], p[2 * ld + i], p[3 * ld + i]};
}
__device__ __forceinline__ void st4(double *p, int ld, int i, dq q) {
  p[i] = q.s; p[ld + i] = q.x; p[2 * ld + i] = q.y; p[3 * ld + i] = q.z;
}
// AoS small records
__device__ __forceinline__ d3 ga3(const double *p) { return {p[0], p[1], p[2]}; }
__device__ __forceinline__ void sa3(double *p, d3 v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }
__device__ __forceinline__ dq ga4(const double *p) { return {p[0], p[1], p[2], p[3]}; }
__device__ __forceinline__ void sa4(double *p, dq q) { p[0] = q.s; p[1] = q.x; p[2] = q.y; p[3] = q.z; }

}  // namespace gmi
