// gmi_pair.hip -- gfx950 RPY pair kernels (all four mobility blocks).
//
// One launch evaluates any subset of the four reference blocks
//   ss: segment (v, w) <- segment (f, tau)   cuda_functions.cu:246-425
//   sb: segment (v, w) <- blob f             cuda_functions.cu:541-694
//   bs: blob v         <- segment (f, tau)   cuda_functions.cu:697-863
//   bb: blob v         <- blob f             cuda_functions.cu:429-539
// over a work list of (target tile, source split) items.  Each workgroup
// owns kBlock targets (one per lane, 4 waves of 64) and walks one contiguous
// slice of its group's sources in LDS tiles of kTile.  Every lane of a wave
// reads the same LDS source at the same time (a broadcast read), so LDS
// bandwidth per pair is one source record per wave.
//
// Far pairs (the common case) use the unnormalised separation d = x_i - x_j,
// powers of 1/|d| from one reciprocal square root, and sources staged
// pre-scaled by the mobility constants (the pair forms are written out above
// pair_ss).  Near pairs (|d| <= a_i + a_j, including the self pair) take a
// wave-uniform branch that reproduces the reference's overlap formulas with
// r_hat = d / (|d| + 1e-20).
//
// FP64 variant: v_rsq_f64 refined by one third-order Newton step.
// FP32 variant: positions split hi + lo so that near separations keep full
// single precision, v_rsq_f32, plain FP32 sums inside a tile and Kahan
// compensated sums across tiles; partial sums across source splits are
// reduced in FP64.
#include "gmi_internal.hpp"
#include <climits>
#include "gmi_wave.hpp"

namespace gmi {
namespace {

constexpr double kPi = 3.14159265358979323846264338327950288;

template <typename R> __device__ __forceinline__ R rsq(R x);
template <> __device__ __forceinline__ float rsq<float>(float x) {
  return __builtin_amdgcn_rsqf(x);
}
template <> __device__ __forceinline__ double rsq<double>(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double e = fma(-(x * y), y, 1.0);             // 1 - x y^2
  return fma(y * e, fma(0.375, e, 0.5), y);          // third-order step
}

// Constants of one radius pair (a_i target, a_j source), precomputed in FP64.
// Sources are staged pre-scaled: F' = f/(8 pi mu), T' = tau/(16 pi mu) and
// T2 = 2 T' (see the pair forms below), so the far-field coefficients carry
// no mobility constants and the overlap coefficients carry 8 pi mu or
// 16 pi mu folded into their constants.
template <typename R> struct PC {
  R far2;       // (a_i + a_j)^2: pairs with r^2 above it use the far forms
  R s3;         // (a_i^2 + a_j^2)/3
  R s;          // a_i^2 + a_j^2
  // near-field (overlap) constants, equal radius (ss / bb)
  R inv_a;      // 1/a
  R c3_32a;     // 3/(32 a)
  R c3_8a;      // 3/(8 a)
  R c6a;        // 8 pi mu / (6 pi mu a)       = 4/(3a)
  R c8a3;       // 16 pi mu / (8 pi mu a^3)    = 2/a^3
  R c16a3;      // 8 pi mu / (16 pi mu a^3)    = 1/(2 a^3)
  // unequal radii (sb / bs): as segment, ab blob
  R as, ab, dlt;
  R c6ab;       // 8 pi mu / (6 pi mu as ab)
  R c128;       // 8 pi mu / (128 pi mu ab as^3)
  R c6max;      // 8 pi mu / (6 pi mu max(as, ab))
  R c8as3;      // 8 pi mu / (8 pi mu as^3) when as > ab, else 0
};

template <typename R>
__host__ __device__ PC<R> make_pc(double ai, double aj) {
  PC<R> k;
  const double sig = ai * ai + aj * aj;
  k.far2 = R((ai + aj) * (ai + aj));
  k.s3 = R(sig / 3.0);
  k.s = R(sig);
  k.inv_a = R(1.0 / ai);
  k.c3_32a = R(3.0 / (32.0 * ai));
  k.c3_8a = R(3.0 / (8.0 * ai));
  k.c6a = R(4.0 / (3.0 * ai));
  k.c8a3 = R(2.0 / (ai * ai * ai));
  k.c16a3 = R(0.5 / (ai * ai * ai));
  return k;
}

// Unequal-radius constants: as = segment radius, ab = blob radius.
template <typename R>
__host__ __device__ PC<R> make_pc_unequal(double as, double ab) {
  PC<R> k = make_pc<R>(as, ab);
  k.as = R(as);
  k.ab = R(ab);
  k.dlt = R(as > ab ? as - ab : ab - as);
  k.c6ab = R(4.0 / (3.0 * as * ab));
  k.c128 = R(1.0 / (16.0 * ab * as * as * as));
  k.c6max = R(4.0 / (3.0 * (as > ab ? as : ab)));
  k.c8as3 = R(as > ab ? 1.0 / (as * as * as) : 0.0);
  return k;
}

// r and 1/r of an overlapping pair from the far field's 1/sqrt(r^2): the
// reference's 1/(r + 1e-20) is 1/r for every r > 0 at this precision, and
// stays finite (1e20) for the self pair, r = 0, where ri is not.
template <typename R>
__device__ __forceinline__ void near_r(R r2, R ri, R &r, R &rm1) {
  const bool pos = r2 > R(0);
  r = pos ? r2 * ri : R(0);
  rm1 = pos ? ri : R(1e20);
}

// Equal-radius overlap coefficients (cuda_functions.cu:336-350, 364-386,
// 400-406) in the scaled pair form: fd = F'.d and td = T'.d.  E is the
// difference of the f x d and tau coefficients of w (zero in the far field).
template <typename R>
__device__ __forceinline__ void near_equal(
    R r2, R ri, R fd, R td, const PC<R> &k, R &At, R &Bt, R &Br, R &Cg, R &E) {
  R r, rm1;
  near_r(r2, ri, r, rm1);
  const R b = r * k.c3_32a;
  At = (R(1) - R(3) * b) * k.c6a;
  Bt = ((b * k.c6a * fd) * rm1) * rm1;
  const R t = r * k.inv_a;
  const R t3 = t * t * t;
  const R Ar = (R(1) - R(27) * t / R(32) + R(5) * t3 / R(64)) * k.c8a3;  // w = Ar T'
  Br = (((R(9) * t / R(32) - R(3) * t3 / R(64)) * k.c8a3 * td) * rm1) * rm1;
  Cg = (r * (R(1) - r * k.c3_8a) * k.c16a3) * rm1;
  E = Cg + Ar;
}

// Unequal-radius overlap coefficients (cuda_functions.cu:622-646, 656-664).
template <typename R>
__device__ __forceinline__ void near_unequal(
    R r2, R ri, R fd, const PC<R> &k, R &At, R &Bt, R &Cg) {
  R r, rm1;
  near_r(r2, ri, r, rm1);
  if (r > k.dlt) {
    const R inv_r3x32 = rm1 * rm1 * rm1 * R(1.0 / 32.0);  // 1/(32 r^3)
    const R p = k.dlt * k.dlt + R(3) * r * r;
    const R q = k.dlt * k.dlt - r * r;
    At = (R(0.5) * (k.as + k.ab) - p * p * inv_r3x32) * k.c6ab;
    Bt = (((R(3) * q * q * inv_r3x32) * k.c6ab * fd) * rm1) * rm1;
    const R s = k.as - k.ab + r;
    const R h = k.as - r;
    Cg = (s * s * rm1 * rm1 *
          (k.ab * k.ab + R(2) * k.ab * (r + k.as) - R(3) * h * h) * k.c128) *
         rm1;
  } else {
    At = k.c6max;
    Bt = R(0);
    Cg = (r * k.c8as3) * rm1;
  }
}

template <typename R> struct Acc {
  R vx, vy, vz, wx, wy, wz;
  __device__ void zero() { vx = vy = vz = wx = wy = wz = R(0); }
};

// One staged source, one record per source so that a pair reads it with
// 16-byte LDS loads from one address (every lane reads the same source, a
// broadcast): position (FP32 keeps a low part too), F', and for segment
// sources T' and T2.  What a translation-only pair needs fills the first 48
// bytes.
template <typename R> struct Src;
template <> struct alignas(16) Src<float> {
  float x, y, z, lx, ly, lz, fx, fy, fz, tx, ty, tz, ux, uy, uz, pad;
};
template <> struct alignas(16) Src<double> {
  double x, y, z, fx, fy, fz, tx, ty, tz, ux, uy, uz;
};
template <typename R> struct Lds {
  Src<R> s[kTile];
};

// A staged source: the first 48 bytes, and all of it when SIX.
template <bool SIX>
__device__ __forceinline__ Src<float> fetch(const Src<float> *src) {
  const float4 *q = reinterpret_cast<const float4 *>(src);
  const float4 a = q[0], b = q[1], c = q[2];
  Src<float> v;
  v.x = a.x; v.y = a.y; v.z = a.z; v.lx = a.w;
  v.ly = b.x; v.lz = b.y; v.fx = b.z; v.fy = b.w;
  v.fz = c.x; v.tx = c.y; v.ty = c.z; v.tz = c.w;
  if (SIX) {
    const float4 d = q[3];
    v.ux = d.x; v.uy = d.y; v.uz = d.z;
  } else {
    v.ux = v.uy = v.uz = 0.f;
  }
  v.pad = 0.f;
  return v;
}
template <bool SIX>
__device__ __forceinline__ Src<double> fetch(const Src<double> *src) {
  const double2 *q = reinterpret_cast<const double2 *>(src);
  const double2 a = q[0], b = q[1], c = q[2];
  Src<double> v;
  v.x = a.x; v.y = a.y; v.z = b.x; v.fx = b.y; v.fy = c.x; v.fz = c.y;
  if (SIX) {
    const double2 d = q[3], e = q[4], f = q[5];
    v.tx = d.x; v.ty = d.y; v.tz = e.x; v.ux = e.y; v.uy = f.x; v.uz = f.y;
  } else {
    v.tx = v.ty = v.tz = v.ux = v.uy = v.uz = 0.0;
  }
  return v;
}

// ---- per-pair kernels ------------------------------------------------------
// With d = x_i - x_j, ri = 1/|d| and the pre-scaled sources, the far forms
// (cuda_functions.cu:302-398, 478-517, 593-667, 755-829) are
//   v += At F' + Bt d + Cg (T2 x d)
//   w += Cg (F' x d - T') + Br d
// with At = ri + s3 ri^3, Bt = (ri^3 - s ri^5) F'.d, Br = 3 ri^5 T'.d and
// Cg = ri^3: one reciprocal square root, and every product folded into the
// running sums as a fused multiply-add.  The overlap forms (r <= a_i + a_j,
// including the self pair) run only when some lane of the wave has an
// overlapping pair (a wave-uniform branch, so the common case needs no
// per-pair execution-mask changes) and are selected per lane; there the f x d
// and tau coefficients of w differ by E, added as E T'.  The far values of an
// overlapping pair may be infinite (r = 0); the select discards them.
__device__ __forceinline__ bool wave_any(bool p) { return __ballot(p) != 0; }

// v += At F' + Bt d
template <typename R>
__device__ __forceinline__ void acc_v(R At, R Bt, R dx, R dy, R dz, const Src<R> &s,
                                      Acc<R> &o) {
  o.vx = fma(At, s.fx, o.vx); o.vx = fma(Bt, dx, o.vx);
  o.vy = fma(At, s.fy, o.vy); o.vy = fma(Bt, dy, o.vy);
  o.vz = fma(At, s.fz, o.vz); o.vz = fma(Bt, dz, o.vz);
}

// v += Cg (T2 x d)
template <typename R>
__device__ __forceinline__ void acc_vrot(R Cg, R dx, R dy, R dz, const Src<R> &s,
                                         Acc<R> &o) {
  const R ux = s.ux, uy = s.uy, uz = s.uz;
  o.vx = fma(Cg, fma(uy, dz, -(uz * dy)), o.vx);
  o.vy = fma(Cg, fma(uz, dx, -(ux * dz)), o.vy);
  o.vz = fma(Cg, fma(ux, dy, -(uy * dx)), o.vz);
}

// ss: 6x6 equal radius.
template <typename R>
__device__ __forceinline__ void pair_ss(R dx, R dy, R dz, const Src<R> &s,
                                        const PC<R> &k, Acc<R> &o) {
  const R fx = s.fx, fy = s.fy, fz = s.fz;
  const R tx = s.tx, ty = s.ty, tz = s.tz;
  const R r2 = fma(dz, dz, fma(dy, dy, dx * dx));
  const R fd = fma(fz, dz, fma(fy, dy, fx * dx));
  const R td = fma(tz, dz, fma(ty, dy, tx * dx));
  const R ri = rsq(r2), ri2 = ri * ri, ri3 = ri2 * ri, ri5 = ri3 * ri2;
  R At = fma(k.s3, ri3, ri);
  R Bt = fma(-k.s, ri5, ri3) * fd;
  R Br = (R(3) * ri5) * td;
  R Cg = ri3;
  const bool nr = !(r2 > k.far2);
  if (__builtin_expect(wave_any(nr), 0)) {
    R At2, Bt2, Br2, Cg2, E2;
    near_equal(r2, ri, fd, td, k, At2, Bt2, Br2, Cg2, E2);
    At = nr ? At2 : At; Bt = nr ? Bt2 : Bt; Br = nr ? Br2 : Br; Cg = nr ? Cg2 : Cg;
    const R E = nr ? E2 : R(0);
    o.wx = fma(E, tx, o.wx); o.wy = fma(E, ty, o.wy); o.wz = fma(E, tz, o.wz);
  }
  acc_vrot(Cg, dx, dy, dz, s, o);
  acc_v(At, Bt, dx, dy, dz, s, o);
  o.wx = fma(Cg, fma(-fz, dy, fma(fy, dz, -tx)), o.wx); o.wx = fma(Br, dx, o.wx);
  o.wy = fma(Cg, fma(-fx, dz, fma(fz, dx, -ty)), o.wy); o.wy = fma(Br, dy, o.wy);
  o.wz = fma(Cg, fma(-fy, dx, fma(fx, dy, -tz)), o.wz); o.wz = fma(Br, dz, o.wz);
}

// sb: segment (v, w) <- blob f, unequal radii.
template <typename R>
__device__ __forceinline__ void pair_sb(R dx, R dy, R dz, const Src<R> &s,
                                        const PC<R> &k, Acc<R> &o) {
  const R fx = s.fx, fy = s.fy, fz = s.fz;
  const R r2 = fma(dz, dz, fma(dy, dy, dx * dx));
  const R fd = fma(fz, dz, fma(fy, dy, fx * dx));
  const R ri = rsq(r2), ri2 = ri * ri, ri3 = ri2 * ri, ri5 = ri3 * ri2;
  R At = fma(k.s3, ri3, ri);
  R Bt = fma(-k.s, ri5, ri3) * fd;
  R Cg = ri3;
  const bool nr = !(r2 > k.far2);
  if (__builtin_expect(wave_any(nr), 0)) {
    R At2, Bt2, Cg2;
    near_unequal(r2, ri, fd, k, At2, Bt2, Cg2);
    At = nr ? At2 : At; Bt = nr ? Bt2 : Bt; Cg = nr ? Cg2 : Cg;
  }
  acc_v(At, Bt, dx, dy, dz, s, o);
  o.wx = fma(Cg, fma(fy, dz, -(fz * dy)), o.wx);
  o.wy = fma(Cg, fma(fz, dx, -(fx * dz)), o.wy);
  o.wz = fma(Cg, fma(fx, dy, -(fy * dx)), o.wz);
}

// bs: blob v <- segment (f, tau), unequal radii.
template <typename R>
__device__ __forceinline__ void pair_bs(R dx, R dy, R dz, const Src<R> &s,
                                        const PC<R> &k, Acc<R> &o) {
  const R fx = s.fx, fy = s.fy, fz = s.fz;
  const R r2 = fma(dz, dz, fma(dy, dy, dx * dx));
  const R fd = fma(fz, dz, fma(fy, dy, fx * dx));
  const R ri = rsq(r2), ri2 = ri * ri, ri3 = ri2 * ri, ri5 = ri3 * ri2;
  R At = fma(k.s3, ri3, ri);
  R Bt = fma(-k.s, ri5, ri3) * fd;
  R Cg = ri3;
  const bool nr = !(r2 > k.far2);
  if (__builtin_expect(wave_any(nr), 0)) {
    R At2, Bt2, Cg2;
    near_unequal(r2, ri, fd, k, At2, Bt2, Cg2);
    At = nr ? At2 : At; Bt = nr ? Bt2 : Bt; Cg = nr ? Cg2 : Cg;
  }
  acc_vrot(Cg, dx, dy, dz, s, o);
  acc_v(At, Bt, dx, dy, dz, s, o);
}

// translation only, unequal radii: segment v <- blob f or blob v <- segment f
// (the tt parts of sb / bs; the prescribed-cilia operator has no torques and
// no angular velocities, cuda_functions.cu:646,808)
template <typename R>
__device__ __forceinline__ void pair_ut(R dx, R dy, R dz, const Src<R> &s,
                                        const PC<R> &k, Acc<R> &o) {
  const R fx = s.fx, fy = s.fy, fz = s.fz;
  const R r2 = fma(dz, dz, fma(dy, dy, dx * dx));
  const R fd = fma(fz, dz, fma(fy, dy, fx * dx));
  const R ri = rsq(r2), ri2 = ri * ri, ri3 = ri2 * ri, ri5 = ri3 * ri2;
  R At = fma(k.s3, ri3, ri);
  R Bt = fma(-k.s, ri5, ri3) * fd;
  const bool nr = !(r2 > k.far2);
  if (__builtin_expect(wave_any(nr), 0)) {
    R At2, Bt2, Cg;
    near_unequal(r2, ri, fd, k, At2, Bt2, Cg);
    At = nr ? At2 : At; Bt = nr ? Bt2 : Bt;
  }
  acc_v(At, Bt, dx, dy, dz, s, o);
}

// bb: blob v <- blob f, equal radius (also segment v <- segment f when only
// translations couple).
template <typename R>
__device__ __forceinline__ void pair_bb(R dx, R dy, R dz, const Src<R> &s,
                                        const PC<R> &k, Acc<R> &o) {
  const R fx = s.fx, fy = s.fy, fz = s.fz;
  const R r2 = fma(dz, dz, fma(dy, dy, dx * dx));
  const R fd = fma(fz, dz, fma(fy, dy, fx * dx));
  const R ri = rsq(r2), ri2 = ri * ri, ri3 = ri2 * ri, ri5 = ri3 * ri2;
  R At = fma(k.s3, ri3, ri);
  R Bt = fma(-k.s, ri5, ri3) * fd;
  const bool nr = !(r2 > k.far2);
  if (__builtin_expect(wave_any(nr), 0)) {
    R At2, Bt2, Br, Cg, E;
    near_equal(r2, ri, fd, R(0), k, At2, Bt2, Br, Cg, E);
    At = nr ? At2 : At; Bt = nr ? Bt2 : Bt;
  }
  acc_v(At, Bt, dx, dy, dz, s, o);
}

// ---- precision traits ----------------------------------------------------
// Position of one particle in the kernel's arithmetic: FP64 keeps the
// double; FP32 keeps hi = float(x) and lo = float(x - hi), and forms
// separations as (hi_i - hi_j) + (lo_i - lo_j).
template <typename R> struct Pos;
template <> struct Pos<double> {
  double x, y, z;
  __device__ void load(double px, double py, double pz) { x = px; y = py; z = pz; }
};
template <> struct Pos<float> {
  float x, y, z, lx, ly, lz;
  __device__ void load(double px, double py, double pz) {
    x = float(px); y = float(py); z = float(pz);
    lx = float(px - double(x)); ly = float(py - double(y)); lz = float(pz - double(z));
  }
};

__device__ __forceinline__ void sep(const Pos<double> &p, const Src<double> &s,
                                    double &dx, double &dy, double &dz) {
  dx = p.x - s.x; dy = p.y - s.y; dz = p.z - s.z;
}
__device__ __forceinline__ void sep(const Pos<float> &p, const Src<float> &s,
                                    float &dx, float &dy, float &dz) {
  dx = (p.x - s.x) + (p.lx - s.lx);
  dy = (p.y - s.y) + (p.ly - s.ly);
  dz = (p.z - s.z) + (p.lz - s.lz);
}

// Source record from FP64 position, F' and T2 (T' = T2 / 2).
__device__ __forceinline__ Src<double> make_src(double px, double py, double pz, double fx,
                                                double fy, double fz, double ux, double uy,
                                                double uz, double) {
  Src<double> v;
  v.x = px; v.y = py; v.z = pz; v.fx = fx; v.fy = fy; v.fz = fz;
  v.tx = 0.5 * ux; v.ty = 0.5 * uy; v.tz = 0.5 * uz; v.ux = ux; v.uy = uy; v.uz = uz;
  return v;
}
__device__ __forceinline__ Src<float> make_src(double px, double py, double pz, double fx,
                                               double fy, double fz, double ux, double uy,
                                               double uz, float) {
  Src<float> v;
  v.x = float(px); v.y = float(py); v.z = float(pz);
  v.lx = float(px - double(v.x)); v.ly = float(py - double(v.y)); v.lz = float(pz - double(v.z));
  v.fx = float(fx); v.fy = float(fy); v.fz = float(fz);
  v.tx = float(0.5 * ux); v.ty = float(0.5 * uy); v.tz = float(0.5 * uz);
  v.ux = float(ux); v.uy = float(uy); v.uz = float(uz);
  v.pad = 0.f;
  return v;
}

// Running sum of one target: FP64 plain, FP32 Kahan-compensated per tile.
template <typename R> struct Sum;
template <> struct Sum<double> {
  Acc<double> s;
  __device__ void zero() { s.zero(); }
  __device__ void add(const Acc<double> &t) {
    s.vx += t.vx; s.vy += t.vy; s.vz += t.vz;
    s.wx += t.wx; s.wy += t.wy; s.wz += t.wz;
  }
  __device__ void get(double *o) const {
    o[0] = s.vx; o[1] = s.vy; o[2] = s.vz; o[3] = s.wx; o[4] = s.wy; o[5] = s.wz;
  }
};
__device__ __forceinline__ void kahan(float &sum, float &c, float x) {
  const float y = x - c;
  const float t = sum + y;
  c = (t - sum) - y;
  sum = t;
}
template <> struct Sum<float> {
  Acc<float> s, c;
  __device__ void zero() { s.zero(); c.zero(); }
  __device__ void add(const Acc<float> &t) {
    kahan(s.vx, c.vx, t.vx); kahan(s.vy, c.vy, t.vy); kahan(s.vz, c.vz, t.vz);
    kahan(s.wx, c.wx, t.wx); kahan(s.wy, c.wy, t.wy); kahan(s.wz, c.wz, t.wz);
  }
  __device__ void get(double *o) const {
    o[0] = double(s.vx) - double(c.vx); o[1] = double(s.vy) - double(c.vy);
    o[2] = double(s.vz) - double(c.vz); o[3] = double(s.wx) - double(c.wx);
    o[4] = double(s.wy) - double(c.wy); o[5] = double(s.wz) - double(c.wz);
  }
};

enum Kind { SS, SB, BS, BB, UT };

// Ends a pair's accumulation where it stands: an empty asm that reads and
// writes the six sums, so the compiler cannot defer the sums of one pair past
// the next pair (which keeps every pair's intermediates live at once).
template <typename R> __device__ __forceinline__ void settle(Acc<R> &o) {
  asm volatile("" : "+v"(o.vx), "+v"(o.vy), "+v"(o.vz), "+v"(o.wx), "+v"(o.wy), "+v"(o.wz));
}

// Target p against one staged source.
template <typename R, int K>
__device__ __forceinline__ void pair_at(const Pos<R> &p, const Src<R> *src,
                                        const PC<R> &k, Acc<R> &acc) {
  const Src<R> s = fetch<K == SS || K == BS>(src);
  R dx, dy, dz;
  sep(p, s, dx, dy, dz);
  if (K == SS)
    pair_ss<R>(dx, dy, dz, s, k, acc);
  else if (K == SB)
    pair_sb<R>(dx, dy, dz, s, k, acc);
  else if (K == BS)
    pair_bs<R>(dx, dy, dz, s, k, acc);
  else if (K == UT)
    pair_ut<R>(dx, dy, dz, s, k, acc);
  else
    pair_bb<R>(dx, dy, dz, s, k, acc);
}

// Walk sources [j0, j1) of one species in LDS tiles and accumulate into sum.
// c8 = 1/(8 pi mu) scales the staged sources (F' = c8 f, T' = c8 tau / 2,
// T2 = c8 tau).
template <typename R, int K>
__device__ __forceinline__ void sweep(Lds<R> &lds, const Pos<R> &p,
                                      const double *__restrict__ X,
                                      const double *__restrict__ F, int ld,
                                      int j0, int j1, const PC<R> &k, double c8,
                                      Sum<R> &sum) {
  constexpr bool six = (K == SS || K == BS);  // segment sources carry tau
  const int tid = threadIdx.x;
  for (int base = j0; base < j1; base += kTile) {
    const int n = min(kTile, j1 - base);
    __syncthreads();
    if (tid < n) {
      const int j = base + tid;
      double ux = 0.0, uy = 0.0, uz = 0.0;
      if (six) {
        ux = c8 * F[3 * ld + j]; uy = c8 * F[4 * ld + j]; uz = c8 * F[5 * ld + j];
      }
      lds.s[tid] = make_src(X[j], X[ld + j], X[2 * ld + j], c8 * F[j], c8 * F[ld + j],
                            c8 * F[2 * ld + j], ux, uy, uz, R(0));
    }
    __syncthreads();
    Acc<R> acc;
    acc.zero();
    // Four sources per trip: one LDS address serves the four, and the ballot
    // inside each pair keeps the compiler from unrolling the loop itself.
    // The source index is held in a vector register (the empty asm), so the
    // four loads take it with immediate offsets instead of a copy each.
    int j = 0;
    for (; j + 4 <= n; j += 4) {
      int jv = j;
      asm volatile("" : "+v"(jv));
      const Src<R> *src = lds.s + jv;
      pair_at<R, K>(p, src, k, acc);
      settle(acc);
      pair_at<R, K>(p, src + 1, k, acc);
      settle(acc);
      pair_at<R, K>(p, src + 2, k, acc);
      settle(acc);
      pair_at<R, K>(p, src + 3, k, acc);
    }
    for (; j < n; ++j) pair_at<R, K>(p, lds.s + j, k, acc);
    sum.add(acc);
  }
}

template <typename R>
__device__ __forceinline__ void pair_body(const PairArgs &a, Lds<R> &lds) {
  const WorkItem w = a.work[blockIdx.x];
  const TileDesc T = a.tiles[w.tile];
  const int tid = threadIdx.x;
  const bool seg_t = (T.species == 0);
  const bool use_s = seg_t ? (a.mask & kSS) : (a.mask & kBS);
  const bool use_b = seg_t ? (a.mask & kSB) : (a.mask & kBB);
  const int ns = use_s ? T.nsseg : 0;
  const int nb = use_b ? T.nsblob : 0;
  const long L = long(ns) + nb;
  const int beg = int(L * w.split / T.nsplit);
  const int end = int(L * (w.split + 1) / T.nsplit);

  // Lanes past the end of the tile mirror the tile's first target.
  const int ti = T.t0 + (tid < T.nt ? tid : 0);
  Pos<R> p;
  if (seg_t)
    p.load(a.Xs[ti], a.Xs[a.lds + ti], a.Xs[2 * a.lds + ti]);
  else
    p.load(a.Xb[ti], a.Xb[a.ldb + ti], a.Xb[2 * a.ldb + ti]);

  Sum<R> sum;
  sum.zero();
  const double c8 = 1.0 / (8.0 * kPi * a.mu);
  const int s0 = beg, s1 = min(end, ns);
  const int b0 = max(beg, ns) - ns, b1 = end - ns;
  if (seg_t) {
    if (s1 > s0) {
      const PC<R> k = make_pc<R>(a.a_s, a.a_s);
      if (a.tt_only)
        sweep<R, BB>(lds, p, a.Xs, a.Fs, a.lds, T.sseg0 + s0, T.sseg0 + s1, k, c8, sum);
      else
        sweep<R, SS>(lds, p, a.Xs, a.Fs, a.lds, T.sseg0 + s0, T.sseg0 + s1, k, c8, sum);
    }
    if (b1 > b0) {
      const PC<R> k = make_pc_unequal<R>(a.a_s, a.a_b);
      if (a.tt_only)
        sweep<R, UT>(lds, p, a.Xb, a.Fb, a.ldb, T.sblob0 + b0, T.sblob0 + b1, k, c8, sum);
      else
        sweep<R, SB>(lds, p, a.Xb, a.Fb, a.ldb, T.sblob0 + b0, T.sblob0 + b1, k, c8, sum);
    }
  } else {
    if (s1 > s0) {
      const PC<R> k = make_pc_unequal<R>(a.a_s, a.a_b);
      if (a.tt_only)
        sweep<R, UT>(lds, p, a.Xs, a.Fs, a.lds, T.sseg0 + s0, T.sseg0 + s1, k, c8, sum);
      else
        sweep<R, BS>(lds, p, a.Xs, a.Fs, a.lds, T.sseg0 + s0, T.sseg0 + s1, k, c8, sum);
    }
    if (b1 > b0) {
      const PC<R> k = make_pc<R>(a.a_b, a.a_b);
      sweep<R, BB>(lds, p, a.Xb, a.Fb, a.ldb, T.sblob0 + b0, T.sblob0 + b1, k, c8, sum);
    }
  }
  const bool live = tid < T.nt;
  const int nc = (seg_t && !a.tt_only) ? 6 : 3;
  double *V = seg_t ? a.Vs : a.Vb;
  const int ld = seg_t ? a.lds : a.ldb;
  if (T.nsplit == 1) {
    if (!live) return;
    double o[6];
    sum.get(o);
    for (int c = 0; c < nc; ++c) {
      double *v = V + c * ld + T.t0 + tid;
      *v = a.accumulate ? *v + o[c] : o[c];
    }
    return;
  }
  if (live) {
    double o[6];
    sum.get(o);
    double *P = a.part + size_t(T.part0 + w.split) * 6 * kBlock + tid;
    for (int c = 0; c < nc; ++c) P[c * kBlock] = o[c];
  }
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_pair(const PairArgs a) {
  __shared__ Lds<R> lds;
  if (a.skip && *a.skip) return;
  pair_body<R>(a, lds);
}

// Precision picked on the device: FP32-Kahan when *a.sel32 == 0, FP64
// otherwise (the relaxed GMRES switches with a device flag, so one launch
// serves both precisions).  The LDS tile is sized for FP64.
__global__ __launch_bounds__(kBlock) void k_pair_dyn(const PairArgs a) {
  __shared__ Lds<double> lds;
  static_assert(sizeof(Lds<float>) <= sizeof(Lds<double>), "LDS tile");
  if (a.skip && *a.skip) return;
  if (*a.sel32 == 0)
    pair_body<float>(a, *reinterpret_cast<Lds<float> *>(&lds));
  else
    pair_body<double>(a, lds);
}

// ---------------------------------------------------------------------------
// Segment-segment mobility above a no-slip plane wall at z = 0 (Swan & Brady,
// as the reference evaluates it: cuda_functions.cu:8-242, its sign
// convention for the rotation-translation wall terms included).  FP64 only.
// Positions are divided by a and the four blocks scaled by 1/(mu a),
// 1/(mu a^2), 1/(mu a^2), 1/(mu a^3) -- the reference's expression when
// RSEG = MU = 1.  Pairs j != i use the far-field RPY form at any separation.
// Reciprocals and reciprocal square roots come from the hardware estimates
// refined by Newton steps (no IEEE division sequences in the pair loop).

// 1/sqrt(x) for normal x > 0: the hardware estimate and two Newton steps
// (each squares the relative error), within a few ulp of the rounded value.
__device__ __forceinline__ double rsqrt_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
  double r = fma(-x * y, y, 1.0);
  y = fma(0.5 * y, r, y);
  r = fma(-x * y, y, 1.0);
  return fma(0.5 * y, r, y);
}

// 1/x for normal x: the hardware estimate and two Newton steps.
__device__ __forceinline__ double rcp_nr(double x) {
  double y = __builtin_amdgcn_rcp(x);
  double e = fma(-x, y, 1.0);
  y = fma(y, e, y);
  e = fma(-x, y, 1.0);
  return fma(y, e, y);
}

struct WallAcc {
  double vt[3], vr[3], wt[3], wr[3];  // v from forces / torques, w from forces / torques
};

__device__ __forceinline__ void wall_self(double zi, const double *f, const double *t,
                                          WallAcc &A) {
  constexpr double FF = 1.0 / (6.0 * kPi), TF = 1.0 / (8.0 * kPi);
  const double hm1 = 1.0 / zi, hm3 = hm1 * hm1 * hm1, hm4 = hm3 * hm1, hm5 = hm4 * hm1;
  const double par = FF - (9.0 * hm1 - 2.0 * hm3 + hm5) * (FF / 16.0);
  const double perp = FF - (9.0 * hm1 - 4.0 * hm3 + hm5) * (FF / 8.0);
  const double c = (3.0 * FF / 32.0) * hm4;
  A.vt[0] += par * f[0]; A.vt[1] += par * f[1]; A.vt[2] += perp * f[2];
  A.vr[0] += c * t[1]; A.vr[1] -= c * t[0];
  const double rpar = TF - (15.0 * FF / 64.0) * hm3, rperp = TF - (3.0 * FF / 32.0) * hm3;
  A.wr[0] += rpar * t[0]; A.wr[1] += rpar * t[1]; A.wr[2] += rperp * t[2];
  A.wt[0] -= c * f[1]; A.wt[1] += c * f[0];
}

__device__ __forceinline__ void wall_pair(double dx, double dy, double dz, double zi, double zj,
                                          const double *f, const double *t, WallAcc &A) {
  constexpr double FF = 1.0 / (6.0 * kPi), TF = 1.0 / (8.0 * kPi);
  // unbounded RPY, far-field form
  double rm1 = rsqrt_nr(dx * dx + dy * dy + dz * dz);
  const double rx = dx * rm1, ry = dy * rm1, rz = dz * rm1;
  const double rf = rx * f[0] + ry * f[1] + rz * f[2];
  const double rt = rx * t[0] + ry * t[1] + rz * t[2];
  double a0 = TF * rm1;
  const double b1 = 1.0 + (2.0 / 3.0) * rm1 * rm1, b2 = 1.0 - 2.0 * rm1 * rm1;
  A.vt[0] += a0 * (b1 * f[0] + b2 * rf * rx);
  A.vt[1] += a0 * (b1 * f[1] + b2 * rf * ry);
  A.vt[2] += a0 * (b1 * f[2] + b2 * rf * rz);
  a0 = 0.5 * TF * rm1 * rm1 * rm1;
  A.wr[0] += a0 * (3.0 * rt * rx - t[0]);
  A.wr[1] += a0 * (3.0 * rt * ry - t[1]);
  A.wr[2] += a0 * (3.0 * rt * rz - t[2]);
  a0 = TF * rm1 * rm1;
  A.vr[0] += a0 * (t[1] * rz - t[2] * ry);
  A.vr[1] += a0 * (t[2] * rx - t[0] * rz);
  A.vr[2] += a0 * (t[0] * ry - t[1] * rx);
  A.wt[0] += a0 * (f[1] * rz - f[2] * ry);
  A.wt[1] += a0 * (f[2] * rx - f[0] * rz);
  A.wt[2] += a0 * (f[0] * ry - f[1] * rx);
  // wall corrections: the image of source j at (xj, yj, -zj)
  const double sz = zi + zj;
  rm1 = rsqrt_nr(dx * dx + dy * dy + sz * sz);
  const double rm2 = rm1 * rm1, rm3 = rm2 * rm1, rm4 = rm3 * rm1, rm5 = rm4 * rm1;
  double ex = dx * rm1, ey = dy * rm1;
  const double ez = sz * rm1, ez2 = ez * ez;
  double h = zj * rcp_nr(sz);
  const double ef = ex * f[0] + ey * f[1] + ez * f[2];
  const double et = ex * t[0] + ey * t[1] + ez * t[2];
  {
    const double c0 = -(0.75 * rm1 * (1.0 + 2.0 * h * (1.0 - h) * ez2) +
                        0.5 * (rm3 * (1.0 - 3.0 * ez2) - rm5 * (1.0 - 5.0 * ez2)));
    const double c1 = -(0.75 * rm1 * (1.0 - 6.0 * h * (1.0 - h) * ez2) -
                        1.5 * rm3 * (1.0 - 5.0 * ez2) + 2.5 * rm5 * (1.0 - 7.0 * ez2)) * ef;
    const double c2 = ez * f[2] * (1.5 * h * rm1 * (1.0 - 6.0 * (1.0 - h) * ez2) -
                                   3.0 * rm3 * (1.0 - 5.0 * ez2) + 5.0 * rm5 * (2.0 - 7.0 * ez2));
    const double c3 = (1.5 * h * rm1 - 5.0 * rm5) * ef;
    const double c4 = -((3.0 * h * h * rm1 + 3.0 * rm3) * ez2 + (2.0 - 15.0 * ez2) * rm5) * f[2];
    A.vt[0] += FF * (c0 * f[0] + (c1 + c2) * ex);
    A.vt[1] += FF * (c0 * f[1] + (c1 + c2) * ey);
    A.vt[2] += FF * (c0 * f[2] + (c1 + c2 + c3) * ez + c4);
  }
  {
    const double c0 = rm3 * (0.375 - 2.25 * ez2);
    const double c1 = -1.125 * rm3 * et;
    const double c2 = 2.25 * ez * rm3 * et;
    const double c3 = 2.25 * rm3 * (ex * t[1] - ey * t[0]);
    A.wr[0] += FF * (c0 * t[0] + c1 * ex - c3 * ey);
    A.wr[1] += FF * (c0 * t[1] + c1 * ey + c3 * ex);
    A.wr[2] += FF * (c0 * t[2] + c1 * ez + c2);
  }
  const double q0 = 0.75 * rm2;
  {
    const double c1 = f[2] * (9.0 * h * ez2 * rm2 + (1.5 - 15.0 * ez2) * rm4);
    const double c2 = -ez * ef * (4.5 * h * rm2 - 7.5 * rm4);
    const double c3 = -1.5 * ez * (h * rm2 - rm4);
    A.wt[0] -= FF * (q0 * (f[1] * ez - f[2] * ey) - (c1 + c2) * ey + c3 * f[1]);
    A.wt[1] -= FF * (q0 * (f[2] * ex - f[0] * ez) + (c1 + c2) * ex - c3 * f[0]);
    A.wt[2] -= FF * q0 * (f[0] * ey - f[1] * ex);
  }
  {  // the rotation-translation block of the pair (j, i), transposed
    h = 1.0 - h;
    ex = -ex;
    ey = -ey;
    const double ect = ex * t[1] - ey * t[0];
    const double c1 = (9.0 * h * ez2 * rm2 + (1.5 - 15.0 * ez2) * rm4) * ect;
    const double c2 = -ez * (4.5 * h * rm2 - 7.5 * rm4) * ect;
    const double c3 = -1.5 * ez * (h * rm2 - rm4);
    A.vr[0] -= FF * (q0 * (t[2] * ey - t[1] * ez) + c2 * ex - c3 * t[1]);
    A.vr[1] -= FF * (q0 * (t[0] * ez - t[2] * ex) + c2 * ey + c3 * t[0]);
    A.vr[2] -= FF * (q0 * (t[1] * ex - t[0] * ey) + c2 * ez + c1);
  }
}

// Translations only (segments carry forces, no torques): the velocity rows
// of wall_self / wall_pair from the force columns, in the same order.
__device__ __forceinline__ void wall_self_tt(double zi, const double *f, double *vt) {
  constexpr double FF = 1.0 / (6.0 * kPi);
  const double hm1 = 1.0 / zi, hm3 = hm1 * hm1 * hm1, hm4 = hm3 * hm1, hm5 = hm4 * hm1;
  const double par = FF - (9.0 * hm1 - 2.0 * hm3 + hm5) * (FF / 16.0);
  const double perp = FF - (9.0 * hm1 - 4.0 * hm3 + hm5) * (FF / 8.0);
  vt[0] += par * f[0]; vt[1] += par * f[1]; vt[2] += perp * f[2];
}

// The same velocity rows with the reference's expressions rewritten
// without divisions: with sz = zi + zj, h = zj / sz and ez = sz / R,
//   h (1 - h) ez^2 = zi zj / R^2,   h ez / R = zj / R^2,
//   (1 - h) ez^2 = zi sz / R^2,     h^2 ez^2 / R = zj^2 / R^3,
// and the unbounded term as (d . f) d / r^2.  Algebraically identical;
// rounding differs from the literal form by a few ulp per pair.
__device__ __forceinline__ void wall_pair_tt(double dx, double dy, double dz, double zi,
                                             double zj, const double *f, double *vt) {
  constexpr double FF = 1.0 / (6.0 * kPi), TF = 1.0 / (8.0 * kPi);
  const double q = dx * dx + dy * dy;
  // unbounded RPY, far form: a0 (b1 f + b2 (d.f) d / r^2)
  const double r1 = rsqrt_nr(fma(dz, dz, q));
  const double r2 = r1 * r1;
  const double a0 = TF * r1;
  const double b1 = a0 * fma(2.0 / 3.0, r2, 1.0);
  const double cd = a0 * fma(-2.0, r2, 1.0) * r2 * (dx * f[0] + dy * f[1] + dz * f[2]);
  vt[0] += b1 * f[0] + cd * dx;
  vt[1] += b1 * f[1] + cd * dy;
  vt[2] += b1 * f[2] + cd * dz;
  // image corrections
  const double sz = zi + zj;
  const double rm1 = rsqrt_nr(fma(sz, sz, q));
  const double rm2 = rm1 * rm1, rm3 = rm2 * rm1, rm5 = rm3 * rm2;
  const double ex = dx * rm1, ey = dy * rm1, ez = sz * rm1, ez2 = ez * ez;
  const double ef = ex * f[0] + ey * f[1] + ez * f[2];
  const double P = zi * zj * rm2;              // h (1 - h) ez^2
  const double g3 = rm3 * (1.0 - 3.0 * ez2), g5 = rm5 * (1.0 - 5.0 * ez2);
  const double w5 = 1.0 - 5.0 * ez2;
  const double c0 = -(0.75 * rm1 * (1.0 + 2.0 * P) + 0.5 * (g3 - g5));
  const double c1 = -(0.75 * rm1 * (1.0 - 6.0 * P) - 1.5 * rm3 * w5 +
                      2.5 * rm5 * (1.0 - 7.0 * ez2)) * ef;
  const double zr = zj * rm2;                  // h ez / R
  const double c2 = f[2] * (1.5 * zr * (1.0 - 6.0 * zi * sz * rm2) +
                            ez * (-3.0 * rm3 * w5 + 5.0 * rm5 * (2.0 - 7.0 * ez2)));
  const double c3ez = (1.5 * zr - 5.0 * rm5 * ez) * ef;
  const double c4 = -(3.0 * zj * zj * rm3 + 3.0 * rm3 * ez2 + (2.0 - 15.0 * ez2) * rm5) * f[2];
  const double c12 = c1 + c2;
  vt[0] += FF * (c0 * f[0] + c12 * ex);
  vt[1] += FF * (c0 * f[1] + c12 * ey);
  vt[2] += FF * (c0 * f[2] + c12 * ez + c3ez + c4);
}

// One lane per target segment, sources staged in LDS tiles.  TT: 3 force
// rows in, 3 velocity rows out (PairArgs::tt_only); else 6 and 6.
template <bool TT>
__global__ __launch_bounds__(kBlock) void k_pair_wall(const PairArgs a) {
  constexpr int NC = TT ? 3 : 6;
  __shared__ double sx[kTile], sy[kTile], sz[kTile], sf[NC][kTile];
  if (a.skip && *a.skip) return;
  const WorkItem w = a.work[blockIdx.x];
  const TileDesc T = a.tiles[w.tile];
  const int tid = threadIdx.x;
  const double inv_a = 1.0 / a.a_s;
  const int ti = T.t0 + (tid < T.nt ? tid : 0);
  const double xi = a.Xs[ti] * inv_a, yi = a.Xs[a.lds + ti] * inv_a, zi = a.Xs[2 * a.lds + ti] * inv_a;
  const int ns = T.nsseg;
  const int beg = int(long(ns) * w.split / T.nsplit);
  const int end = int(long(ns) * (w.split + 1) / T.nsplit);
  WallAcc A;
  for (int k = 0; k < 3; k++) A.vt[k] = A.vr[k] = A.wt[k] = A.wr[k] = 0.0;
  for (int base = T.sseg0 + beg; base < T.sseg0 + end; base += kTile) {
    const int n = min(kTile, T.sseg0 + end - base);
    __syncthreads();
    if (tid < n) {
      const int j = base + tid;
      sx[tid] = a.Xs[j] * inv_a; sy[tid] = a.Xs[a.lds + j] * inv_a; sz[tid] = a.Xs[2 * a.lds + j] * inv_a;
      for (int c = 0; c < NC; c++) sf[c][tid] = a.Fs[c * a.lds + j];
    }
    __syncthreads();
    for (int jj = 0; jj < n; jj++) {
      const double f[3] = {sf[0][jj], sf[1][jj], sf[2][jj]};
      if constexpr (TT) {
        if (base + jj == ti)
          wall_self_tt(zi, f, A.vt);
        else
          wall_pair_tt(xi - sx[jj], yi - sy[jj], zi - sz[jj], zi, sz[jj], f, A.vt);
      } else {
        const double t[3] = {sf[3][jj], sf[4][jj], sf[5][jj]};
        if (base + jj == ti)
          wall_self(zi, f, t, A);
        else
          wall_pair(xi - sx[jj], yi - sy[jj], zi - sz[jj], zi, sz[jj], f, t, A);
      }
    }
  }
  if (tid >= T.nt) return;
  const double s1 = 1.0 / (a.mu * a.a_s), s2 = s1 * inv_a, s3 = s2 * inv_a;
  double o[6];
  for (int k = 0; k < 3; k++) {
    o[k] = s1 * A.vt[k] + s2 * A.vr[k];
    o[3 + k] = s2 * A.wt[k] + s3 * A.wr[k];
  }
  if (T.nsplit == 1) {
    for (int c = 0; c < NC; ++c) {
      double *v = a.Vs + c * a.lds + T.t0 + tid;
      *v = a.accumulate ? *v + o[c] : o[c];
    }
  } else {
    double *P = a.part + size_t(T.part0 + w.split) * 6 * kBlock + tid;
    for (int c = 0; c < NC; ++c) P[c * kBlock] = o[c];
  }
}

// Split partials of one (tile, component) per workgroup (blockIdx.y is the
// component), four lanes per target: lane group g sums the splits
// q = g (mod 4) below the last multiple of four, group 0 also the rest, and
// the four sums are combined in a fixed order -- the sums a single lane with
// four running sums would form, with four times the loads in flight.
constexpr int kReduceGroups = 4;
__global__ __launch_bounds__(kReduceGroups * kBlock) void k_reduce(const PairArgs a) {
  if (a.skip && *a.skip) return;
  __shared__ double sh[kReduceGroups][kBlock];
  const TileDesc T = a.tiles[blockIdx.x];
  const int tid = threadIdx.x & (kBlock - 1), g = threadIdx.x / kBlock, c = blockIdx.y;
  if (T.nsplit == 1) return;
  const bool seg_t = (T.species == 0);
  const int nc = (seg_t && !a.tt_only) ? 6 : 3;
  if (c >= nc) return;
  const bool live = tid < T.nt;
  double s = 0.0;
  if (live) {
    const double *P = a.part + size_t(T.part0) * 6 * kBlock + c * kBlock + tid;
    const size_t stride = size_t(6) * kBlock;
    const int main = T.nsplit & ~(kReduceGroups - 1);
    for (int q = g; q < main; q += kReduceGroups) s += P[stride * q];
    if (g == 0)
      for (int q = main; q < T.nsplit; ++q) s += P[stride * q];
  }
  sh[g][tid] = s;
  __syncthreads();
  if (g != 0 || !live) return;
  const double r = (sh[0][tid] + sh[1][tid]) + (sh[2][tid] + sh[3][tid]);
  double *V = seg_t ? a.Vs : a.Vb;
  const int ld = seg_t ? a.lds : a.ldb;
  double *v = V + c * ld + T.t0 + tid;
  *v = a.accumulate ? *v + r : r;
}

// Steric barrier (cuda_functions.cu:961-1157) over a hashed cell grid.
// Cells are cubes of side h >= 1.1 (a_i + a_j) for every radius pair, so a
// contributing pair lies in the same or adjacent cells.  Each call bins every
// particle (segments first, then blobs, in one index space) by its cell and
// swimmer group into H buckets (a counting sort: integer atomics for the
// counts, then each bucket put in particle-index order), and each target visits
// the 27 cells around its own in a fixed order, skipping entries of other
// cells or groups that share a bucket.  Sums therefore run in a fixed order
// (cell by cell of the 27, index order within a cell).  The contact test compares
// squared distances, so the square root and the division run only for the
// pairs inside 1.1 (a_i + a_j).
__device__ __forceinline__ void barrier_pair(double dx, double dy, double dz, double asum,
                                             double K, double &fx, double &fy, double &fz) {
  const double d2 = dx * dx + dy * dy + dz * dz;
  const double lim = 1.1 * asum;
  if (__builtin_expect(d2 < lim * lim, 0)) {
    const double dist = sqrt(d2);
    double fac = fmin(1.0, 1.0 - (10.0 / asum) * (dist - asum));
    fac = K * fac * fac * fac * fac;
    const double dm1 = 1.0 / dist;
    fx += fac * dx * dm1; fy += fac * dy * dm1; fz += fac * dz * dm1;
  }
}

__device__ __forceinline__ int cell_bucket(int cx, int cy, int cz, int g, int H) {
  const unsigned h = (unsigned(cx) * 73856093u) ^ (unsigned(cy) * 19349663u) ^
                     (unsigned(cz) * 83492791u) ^ (unsigned(g) * 2654435761u);
  return int(h & unsigned(H - 1));
}

__device__ __forceinline__ void particle_pos(int p, const double *Xs, int lds, const double *Xb,
                                             int ldb, int Ns, double &x, double &y, double &z) {
  if (p < Ns) {
    x = Xs[p]; y = Xs[lds + p]; z = Xs[2 * lds + p];
  } else {
    const int q = p - Ns;
    x = Xb[q]; y = Xb[ldb + q]; z = Xb[2 * ldb + q];
  }
}

// cell and group of every particle; bucket counts
__global__ __launch_bounds__(256) void k_cell_key(CellGrid G, const double *Xs, int lds,
                                                  const double *Xb, int ldb, int Ns, int N) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= N) return;
  double x, y, z;
  particle_pos(p, Xs, lds, Xb, ldb, Ns, x, y, z);
  const int4 k = make_int4(int(floor(x * G.inv_h)), int(floor(y * G.inv_h)),
                           int(floor(z * G.inv_h)), G.grp[p]);
  G.key[p] = k;
  const int b = cell_bucket(k.x, k.y, k.z, k.w, G.H);
  G.bucket[p] = b;
  atomicAdd(&G.cnt[b], 1);
}

// exclusive scan of kScanChunk bucket counts per workgroup; chunk totals to boff
__global__ __launch_bounds__(1024) void k_cell_scan_local(CellGrid G) {
  static_assert(kScanChunk == 4 * 1024, "four buckets per thread");
  __shared__ int sh[1024];
  const int t = threadIdx.x;
  const int base = blockIdx.x * kScanChunk + 4 * t;
  int v[4], s = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[k] = G.cnt[base + k];
    s += v[k];
  }
  sh[t] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int add = t >= off ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += add;
    __syncthreads();
  }
  int e = sh[t] - s;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    G.start[base + k] = e;
    e += v[k];
  }
  if (t == 1023) G.boff[blockIdx.x] = sh[1023];
}

// exclusive scan of the chunk totals (at most a few hundred; one thread)
__global__ void k_cell_scan_top(CellGrid G) {
  int acc = 0;
  for (int b = 0; b < G.H / kScanChunk; ++b) {
    const int v = G.boff[b];
    G.boff[b] = acc;
    acc += v;
  }
}

__device__ __forceinline__ int bucket_begin(const CellGrid &G, int b) {
  return G.start[b] + G.boff[b / kScanChunk];
}
__device__ __forceinline__ int bucket_end(const CellGrid &G, int b, int N) {
  return b + 1 < G.H ? bucket_begin(G, b + 1) : N;
}

// every particle into its bucket; the counts are not needed past the scan
// and are cleared for the next build
__global__ __launch_bounds__(256) void k_cell_scatter(CellGrid G, int N) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= N) return;
  const int b = G.bucket[p];
  G.arr[bucket_begin(G, b) + atomicAdd(&G.fill[b], 1)] = p;
  G.cnt[b] = 0;
}

// each particle's place in its bucket: after the scatter a bucket's
// entries (arr) are in arrival order, and a particle's rank among them is the
// number of entries with a smaller index, so the bucket's records (index,
// position, cell and group) land in index order without a sort.  Every
// particle of a bucket clears the bucket's fill counter for the next build
// (all write the same zero, after the scatter's increments).
__global__ __launch_bounds__(256) void k_cell_rank(CellGrid G, const double *Xs, int lds,
                                                   const double *Xb, int ldb, int Ns, int N) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= N) return;
  const int b = G.bucket[p];
  const int s = bucket_begin(G, b), e = bucket_end(G, b, N);
  int r = 0;
  for (int i = s; i < e; ++i) r += G.arr[i] < p;
  const int o = s + r;
  G.ent[o] = p;
  particle_pos(p, Xs, lds, Xb, ldb, Ns, G.rx[o], G.ry[o], G.rz[o]);
  G.rk[o] = G.key[p];
  G.fill[b] = 0;
}

// 27 threads per target particle, one per cell of the 3 x 3 x 3 block around
// its own (thread t of a workgroup takes cell t / kCellTargets for target
// t % kCellTargets); the cells' forces are summed in cell order.  Segments
// add the barrier force to their device force (translation rows), blobs
// write their repulsion force.
constexpr int kCellTargets = 32;
__global__ __launch_bounds__(27 * kCellTargets) void k_barrier_cells(
    CellGrid G, double *__restrict__ Fs, int lds, double *__restrict__ Frep, int ldb,
    const double *__restrict__ Xs, const double *__restrict__ Xb, int Ns, int N, int nseg_fil,
    double as, double ab, double K) {
  __shared__ double sf[27][3][kCellTargets];
  const int lane = threadIdx.x % kCellTargets, cell = threadIdx.x / kCellTargets;
  const int i0 = blockIdx.x * kCellTargets + lane;
  const bool live = i0 < N;
  const int i = live ? i0 : N - 1;
  const bool seg_i = i < Ns;
  double xi, yi, zi;
  particle_pos(i, Xs, lds, Xb, ldb, Ns, xi, yi, zi);
  const int4 ki = G.key[i];
  // excluded sources: a segment itself and its neighbours in the same
  // filament (:1040); a blob itself (:1120)
  int lo = i, hi = i;
  if (seg_i) {
    if (i % nseg_fil != 0) lo = i - 1;
    if ((i + 1) % nseg_fil != 0) hi = i + 1;
  }
  const double ai = seg_i ? as : ab;
  const int cx = ki.x + cell % 3 - 1, cy = ki.y + (cell / 3) % 3 - 1, cz = ki.z + cell / 9 - 1;
  const int b = cell_bucket(cx, cy, cz, ki.w, G.H);
  const int e1 = bucket_end(G, b, N);
  double fx = 0.0, fy = 0.0, fz = 0.0;
  for (int e = bucket_begin(G, b); e < e1; ++e) {
    const int4 kj = G.rk[e];
    if (kj.x != cx || kj.y != cy || kj.z != cz || kj.w != ki.w) continue;
    const int j = G.ent[e];
    const bool seg_j = j < Ns;
    if (seg_i == seg_j && j >= lo && j <= hi) continue;
    barrier_pair(xi - G.rx[e], yi - G.ry[e], zi - G.rz[e], ai + (seg_j ? as : ab), K, fx, fy, fz);
  }
  sf[cell][0][lane] = fx;
  sf[cell][1][lane] = fy;
  sf[cell][2][lane] = fz;
  __syncthreads();
  if (cell != 0 || !live) return;
  double f[3];
  for (int c = 0; c < 3; ++c) {
    double t = 0.0;
    for (int r = 0; r < 27; ++r) t += sf[r][c][lane];
    f[c] = t;
  }
  if (seg_i) {
    Fs[i] += f[0]; Fs[lds + i] += f[1]; Fs[2 * lds + i] += f[2];
  } else {
    const int q = i - Ns;
    Frep[q] = f[0]; Frep[ldb + q] = f[1]; Frep[2 * ldb + q] = f[2];
  }
}

__global__ void k_aos_to_soa(double *__restrict__ dst, int ld,
                             const double *__restrict__ src, int n, int nc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int c = 0; c < nc; ++c) dst[c * ld + i] = src[size_t(nc) * i + c];
}

__global__ void k_soa_to_aos(double *__restrict__ dst,
                             const double *__restrict__ src, int ld, int n,
                             int nc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int c = 0; c < nc; ++c) dst[size_t(nc) * i + c] = src[c * ld + i];
}

}  // namespace

hipError_t launch_pair(const PairArgs &a, int nwork, int ntiles_reduce, bool fp32,
                       hipStream_t s, hipEvent_t ev_mid) {
  if (nwork <= 0) return hipSuccess;
  if (fp32)
    hipLaunchKernelGGL(k_pair<float>, dim3(nwork), dim3(kBlock), 0, s, a);
  else
    hipLaunchKernelGGL(k_pair<double>, dim3(nwork), dim3(kBlock), 0, s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (ev_mid && (e = hipEventRecord(ev_mid, s)) != hipSuccess) return e;
  if (ntiles_reduce > 0) {
    hipLaunchKernelGGL(k_reduce, dim3(ntiles_reduce, 6), dim3(kReduceGroups * kBlock), 0, s, a);
    e = hipGetLastError();
  }
  return e;
}

hipError_t launch_pair_dyn(const PairArgs &a, int nwork, int ntiles_reduce, hipStream_t s) {
  if (nwork <= 0) return hipSuccess;
  if (!a.sel32) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_pair_dyn, dim3(nwork), dim3(kBlock), 0, s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || ntiles_reduce <= 0) return e;
  hipLaunchKernelGGL(k_reduce, dim3(ntiles_reduce, 6), dim3(kReduceGroups * kBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_pair_wall(const PairArgs &a, int nwork, int ntiles_reduce, hipStream_t s,
                            hipEvent_t ev_mid) {
  if (nwork <= 0) return hipSuccess;
  if (a.tt_only)
    hipLaunchKernelGGL(k_pair_wall<true>, dim3(nwork), dim3(kBlock), 0, s, a);
  else
    hipLaunchKernelGGL(k_pair_wall<false>, dim3(nwork), dim3(kBlock), 0, s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (ev_mid && (e = hipEventRecord(ev_mid, s)) != hipSuccess) return e;
  if (ntiles_reduce > 0) {
    hipLaunchKernelGGL(k_reduce, dim3(ntiles_reduce, 6), dim3(kReduceGroups * kBlock), 0, s, a);
    e = hipGetLastError();
  }
  return e;
}

hipError_t launch_barrier(double *Fs, int lds, double *Frep, int ldb, const double *Xs,
                          const double *Xb, int Ns, int Nb, const CellGrid &G, int nseg_fil,
                          double as, double ab, double K, hipStream_t s) {
  const int N = Ns + Nb;
  if (N <= 0) return hipSuccess;
  const int nb = (N + 255) / 256;
  hipLaunchKernelGGL(k_cell_key, dim3(nb), dim3(256), 0, s, G, Xs, lds, Xb, ldb, Ns, N);
  hipLaunchKernelGGL(k_cell_scan_local, dim3(G.H / kScanChunk), dim3(1024), 0, s, G);
  hipLaunchKernelGGL(k_cell_scan_top, dim3(1), dim3(1), 0, s, G);
  hipLaunchKernelGGL(k_cell_scatter, dim3(nb), dim3(256), 0, s, G, N);
  hipLaunchKernelGGL(k_cell_rank, dim3(nb), dim3(256), 0, s, G, Xs, lds, Xb, ldb, Ns, N);
  hipLaunchKernelGGL(k_barrier_cells, dim3((N + kCellTargets - 1) / kCellTargets),
                     dim3(27 * kCellTargets), 0, s, G, Fs, lds, Frep, ldb, Xs, Xb, Ns, N, nseg_fil,
                     as, ab, K);
  return hipGetLastError();
}

hipError_t launch_aos_to_soa(double *dst, int ld, const double *src, int n,
                             int nc, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_aos_to_soa, dim3((n + 255) / 256), dim3(256), 0, s,
                     dst, ld, src, n, nc);
  return hipGetLastError();
}

hipError_t launch_soa_to_aos(double *dst, const double *src, int ld, int n,
                             int nc, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_soa_to_aos, dim3((n + 255) / 256), dim3(256), 0, s,
                     dst, src, ld, n, nc);
  return hipGetLastError();
}

}  // namespace gmi
