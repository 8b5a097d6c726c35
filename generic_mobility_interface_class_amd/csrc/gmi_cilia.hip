// gmi_cilia.hip -- prescribed-cilia time step on the device (include/gmi_cilia.h).
//
// The reference's CILIA_TYPE 3 build with its GMRES solver: filament shapes
// follow the phase (filament.cpp:465-893), and every step solves
//
//   [ M_ss  M_sb  -K_s  -k ] [f_s]   [rhs_s]     K: rigid-body motion
//   [ M_bs  M_bb  -K_b   0 ] [f_b] = [rhs_b]        V + W x d
//   [-K_s^T -K_b^T  0    0 ] [ U ]   [  0  ]     k: velocity directions
//   [ -k^T   0      0    0 ] [phi]   [ -q  ]        dx/dphase per filament
//
// (mobility_solver.cpp:733-904) for segment and blob forces, body velocities
// U = (V, W) and phase speeds, right-preconditioned by the block inverse
// that replaces M by its self-mobility diagonal (:427-650), then moves bodies
// and phases explicitly (cilia_sim_main.cpp:218-232, filament.cpp:1020-1031).
// Free bodies and dynamic phases are each optional (the U and phi rows
// vanish when the body velocity or the phase speed is prescribed).  Above a
// no-slip wall (the context's wall flag) the carpet has no blobs and no body
// rows, and M_ss is the translation part of the wall kernel.
//
// Device layout of a system vector: segment forces SoA [3][lds] | blob forces
// SoA [3][ldb] | U [6 Nsw] | phi [Nf], zero padding -- the pair kernels read
// the force blocks in place.  Decomposition:
//   per filament : frame, shape, velocity directions  (workgroup of 64)
//   per swimmer  : (K^T M0^-1 K)^-1 with the phase     (workgroup of 256)
//                  Schur term (swimmer.cpp:929-1027)
//   per tile     : preconditioner and K products over  (workgroup of 256; a
//                  up to 256 blobs or whole filaments    tile never splits a
//                  of one swimmer                        filament)
//   per vector   : GMRES (gmi_krylov.hpp)
#include "gmi_internal.hpp"
#include "gmi_ctx.hpp"
#include "gmi_lie.hpp"
#include "gmi_krylov.hpp"
#include "../../include/gmi_cilia.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace gmi {
namespace {

constexpr double kPiC = 3.14159265358979323846264338327950288;
constexpr int kMaxSegPerFil = 256;

// body record per swimmer (doubles)
enum { CB_X = 0, CB_Q = 3, CB_QM1 = 7, CB_U = 11, CB_VP = 14, CB_KI = 20, CB_REC = 56 };
// per-filament vectors: |k|^2, sum k, sum (x - X) x k
enum { FV_KN = 0, FV_V1 = 1, FV_V2 = 4, FV_REC = 8 };

// Beat of the filaments.  Fitted beats: Fourier coefficients, rows = cosine
// modes 0..3 (A) or sine modes 1..3 (B), columns = powers s, s^2, s^3
// (filament.cpp:471-549, the reference's transposed storage with the zero
// modes halved).  Build-a-beat: the stroke windows (filament.cpp:719-763).
struct Beat {
  int kind;
  double theta, wR, wE, zva;
  double Ax[12], Ay[12], Bx[9], By[9];
};

// filament.cpp:471-520 (Ito, Omori & Ishikawa 2019, Table 1, rescaled)
constexpr double kFbAy[12] = {-3.3547e-01, 4.0369e-01, 1.0362e-01, 4.0318e-01,
                              -1.5553e+00, 7.3455e-01, -9.9513e-02, 3.2829e-02,
                              -1.2106e-01, 8.1046e-02, -3.0982e-01, 1.4568e-01};
constexpr double kFbAx[12] = {9.7204e-01, -2.8315e-01, 4.9243e-02, -1.8466e-02,
                              -1.2926e-01, 2.6981e-01, 1.6209e-01, -3.4983e-01,
                              1.9082e-01, 1.0259e-02, 3.5907e-02, -6.8736e-02};
constexpr double kFbBy[9] = {2.9136e-01, 1.0721e+00, -1.0433e+00, 6.1554e-03, 3.2521e-01,
                             -2.8315e-01, -6.0528e-02, 2.3185e-01, -2.0108e-01};
constexpr double kFbBx[9] = {1.9697e-01, -5.1193e-01, 3.4778e-01, -5.1295e-02, 4.3396e-01,
                             -3.3547e-01, 1.2311e-02, 1.4157e-01, -1.1695e-01};
// filament.cpp:522-549
constexpr double kClAy[12] = {-2.766076e-01, 6.368662e-01, -8.513947e-02, 3.091205e-01,
                              -1.993712e+00, 1.091847e+00, -2.952120e-01, 7.189021e-01,
                              -6.328589e-01, -9.868018e-02, 3.312812e-01, -2.578206e-01};
constexpr double kClAx[12] = {9.034276e-01, -6.634681e-01, 2.016123e-01, -2.468656e-01,
                              1.215869e+00, -5.867236e-01, 3.075836e-01, -6.966322e-01,
                              4.554729e-01, 2.378489e-02, -1.355582e-01, 7.315831e-02};
constexpr double kClBy[9] = {9.165546e-01, 8.848315e-02, -4.838832e-01, -2.946211e-01,
                             1.057964e+00, -6.335721e-01, 6.122744e-02, -2.319094e-01,
                             1.906292e-01};
constexpr double kClBx[9] = {1.336807e-01, -6.807852e-01, 6.008592e-01, 8.146824e-02,
                             3.472676e-01, -2.744220e-01, 3.615272e-02, 8.619119e-02,
                             -6.122992e-02};

struct CilDev {
  int S, Nf, Nsw, Ns, Nb, lds, ldb, n;  // n: device vector length
  int ob, oc, op;                       // offsets of the blob, U and phi blocks
  int free_body, dyn;
  double mu, as, ab, dt, DL;
  const int *fil_sw, *blob_sw;          // owning swimmer
  double *X, *Xb;                       // positions SoA (the context's Xs, Xb)
  double *Db;                           // R(q) blob_ref, SoA [3][ldb]
  double *Kd;                           // velocity directions SoA [3][lds]
  const double *bref;                   // blob references SoA [3][ldb]
  const double *fref;                   // filament base references [Nf][3]
  double *phase, *omega0, *speed;       // [Nf]
  double *fq, *fqm1;                    // filament frames [Nf][4]
  double *fvec;                         // [Nf][FV_REC]
  double *body;                         // [Nsw][CB_REC]
  const double *gref;                   // generalised-force table
  int ng;
  Beat beat;
};

// up to 256 blobs, or whole filaments, of one swimmer
struct CTile {
  int sw, kind;  // kind 0: segments of filaments f0..f0+nf-1, 1: blobs
  int m0, n;     // first particle, count
  int f0, nf;
};

__device__ __forceinline__ double wsum(double v) { return wave_sum(v); }

// block-wide sums of NV values per thread (256 threads); result in out[] of
// every thread.  sh: 4*NV doubles.
template <int NV>
__device__ void block_sums(const double *v, double *out, double *sh) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int k = 0; k < NV; k++) {
    const double s = wsum(v[k]);
    if (lane == 0) sh[4 * k + w] = s;
  }
  __syncthreads();
  for (int k = 0; k < NV; k++) out[k] = ((sh[4 * k] + sh[4 * k + 1]) + sh[4 * k + 2]) + sh[4 * k + 3];
  __syncthreads();
}

// ---- beat shapes -----------------------------------------------------------

__device__ double fourier(const double *A, const double *B, const double *c, const double *s,
                          const double *w) {
  double out = 0.0;
  for (int j = 0; j < 3; j++) {
    double r = 0.0;
    for (int k = 0; k < 4; k++) r += c[k] * A[3 * k + j];
    double t = 0.0;
    for (int k = 0; k < 3; k++) t += s[k] * B[3 * k + j];
    out += (r + t) * w[j];
  }
  return out;
}

// |tangent| of the fitted shape at s (filament.cpp:557-580); c, s: the
// phase's cosines (1, cos p, cos 2p, cos 3p) and sines (sin p, .., sin 3p)
__device__ double fit_tan_norm(const Beat &b, const double *c, const double *s, double sv) {
  const double w[3] = {1.0, 2.0 * sv, 3.0 * sv * sv};
  const double tx = fourier(b.Ax, b.Bx, c, s, w), ty = fourier(b.Ay, b.By, c, s, w);
  return sqrt(tx * tx + ty * ty);
}

// build-a-beat tangent angle and its phase derivative factor
// (filament.cpp:719-813)
__device__ void bab(const Beat &b, double phase, double sv, double *angle, double *dd) {
  const double phi_tr = b.wE + sv * (2.0 * kPiC - b.wE - b.wR);
  double sp = phase - sv * 2.0 * kPiC * b.zva;
  sp -= 2.0 * kPiC * floor(0.5 * sp / kPiC);
  double delta, d;
  if (sp < b.wE) {
    const double t = kPiC * (b.wE - sp) / b.wE;
    double sn, cs;
    sincos(t, &sn, &cs);
    delta = t - sn * cs;
    d = sn;
    d *= -d / b.wE;
  } else if (sp < phi_tr) {
    delta = 0.0;
    d = 0.0;
  } else if (sp < phi_tr + b.wR) {
    const double t = kPiC * (sp - phi_tr) / b.wR;
    double sn, cs;
    sincos(t, &sn, &cs);
    delta = t - sn * cs;
    d = sn;
    d *= d / b.wR;
  } else {
    delta = kPiC;
    d = 0.0;
  }
  *angle = b.theta + (kPiC - 2.0 * b.theta) * delta / kPiC;
  *dd = d;
}

__device__ __forceinline__ d3 mvec(const m3 &A, d3 v) {
  return {A.a[0] * v.x + A.a[1] * v.y + A.a[2] * v.z, A.a[3] * v.x + A.a[4] * v.y + A.a[5] * v.z,
          A.a[6] * v.x + A.a[7] * v.y + A.a[8] * v.z};
}

// swimmer::initial_guess, body part (swimmer.cpp:242-250): q_{n-1} = q
__global__ void k_cb_guess(CilDev a) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= a.Nsw) return;
  double *b = a.body + size_t(n) * CB_REC;
  for (int k = 0; k < 4; k++) b[CB_QM1 + k] = b[CB_Q + k];
}

// blob positions x + R(q) ref (rigid_body.cpp:334-352) and the rotated
// references R(q) ref used by the K products
__global__ void k_cb_blobs(CilDev a) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= a.Nb) return;
  const double *b = a.body + size_t(a.blob_sw[m]) * CB_REC;
  const m3 R = rotm(ga4(b + CB_Q));
  const d3 d = mvec(R, ld3(a.bref, a.ldb, m));
  st3(a.Db, a.ldb, m, d);
  st3(a.Xb, a.ldb, m, ga3(b + CB_X) + d);
}

// filament::initial_guess, prescribed branch (filament.cpp:825-893): the
// frame follows the body's last rotation, segment 0 sits at the base, the
// other segments and the velocity directions follow the beat at the
// filament's phase.  Then |k|^2, sum k and sum (x - X) x k
// (swimmer.cpp:974-1008).  One workgroup of 64 per filament.
constexpr int kShapeThreads = 64;
__global__ __launch_bounds__(kShapeThreads) void k_cf_shape(CilDev a) {
  __shared__ double mesh[10 * kMaxSegPerFil + 1];   // |tangent| on the length mesh
  __shared__ double cum[10 * kMaxSegPerFil + 1];    // trapezoid partial sums
  __shared__ double su[kMaxSegPerFil];              // arc-length parameters
  __shared__ double sR[9], sbase[3], sL;
  const int f = blockIdx.x, t = threadIdx.x, S = a.S;
  const double *b = a.body + size_t(a.fil_sw[f]) * CB_REC;
  const double phase = a.phase[f];
  if (t == 0) {
    const dq q = qmul(lie_exp(ga3(b + CB_U)), ga4(a.fqm1 + 4 * f));
    sa4(a.fq + 4 * f, q);
    const m3 R = rotm(q);
    for (int k = 0; k < 9; k++) sR[k] = R.a[k];
    const d3 base = ga3(b + CB_X) + mvec(rotm(ga4(b + CB_Q)), ga3(a.fref + 3 * f));
    sbase[0] = base.x; sbase[1] = base.y; sbase[2] = base.z;
    a.speed[f] = a.omega0[f];
  }
  __syncthreads();
  m3 R;
  for (int k = 0; k < 9; k++) R.a[k] = sR[k];
  const d3 base = {sbase[0], sbase[1], sbase[2]};
  const int g0 = f * S;
  const Beat &bt = a.beat;
  if (bt.kind == 0) {
    if (t == 0) {
      double ang, dd;
      bab(bt, phase, 0.0, &ang, &dd);
      const double fac = 2.0 * (kPiC - 2.0 * bt.theta);
      double sn, cs;
      sincos(ang, &sn, &cs);
      d3 t1 = mvec(R, d3{sn, cs, 0.0});
      d3 k1 = mvec(R, d3{fac * dd * cs, -fac * dd * sn, 0.0});
      d3 x = base, kv = {0.0, 0.0, 0.0};
      st3(a.X, a.lds, g0, x);
      st3(a.Kd, a.lds, g0, kv);
      for (int n = 1; n < S; n++) {
        const double sv = double(n) / double(S - 1);
        bab(bt, phase, sv, &ang, &dd);
        sincos(ang, &sn, &cs);
        const d3 t2 = mvec(R, d3{sn, cs, 0.0});
        x = x + (0.5 * a.DL) * (t1 + t2);
        t1 = t2;
        const d3 k2 = mvec(R, d3{fac * dd * cs, -fac * dd * sn, 0.0});
        kv = kv + (0.5 * a.DL) * (k1 + k2);
        k1 = k2;
        st3(a.X, a.lds, g0 + n, x);
        st3(a.Kd, a.lds, g0 + n, kv);
      }
    }
  } else {
    double c[4], s[3];
    c[0] = 1.0;
    for (int n = 1; n < 4; n++) sincos(n * phase, &s[n - 1], &c[n]);
    // filament.cpp:618-666: trapezoids on a fixed mesh of 10 NSEG intervals
    const int ntr = 10 * S;
    const double dl = 1.0 / double(ntr);
    for (int k = t; k <= ntr; k += kShapeThreads) mesh[k] = fit_tan_norm(bt, c, s, k * dl);
    __syncthreads();
    if (t == 0) {
      cum[0] = 0.0;
      for (int k = 1; k <= ntr; k++) cum[k] = cum[k - 1] + 0.5 * dl * (mesh[k - 1] + mesh[k]);
      auto length = [&](double sv) -> double {
        if (!(sv > 0)) return 0.0;
        const double fl = floor(sv / dl), cl = ceil(sv / dl);
        const int ifl = int(fl);
        double len = cum[ifl];
        if (cl > fl) {
          const int icl = int(cl);
          const double f1 = icl <= ntr ? mesh[icl] : fit_tan_norm(bt, c, s, cl * dl);
          len += ((sv / dl - fl) / (cl - fl)) * 0.5 * dl * (mesh[ifl] + f1);
        }
        return len;
      };
      // filament.cpp:668-715: bisection for equal arc-length fractions
      su[0] = 0.0;
      su[S - 1] = 1.0;
      const double total = length(1.0);
      for (int n = 1; n < S - 1; n++) {
        const double target = double(n) / double(S - 1);
        double lo = su[n - 1], hi = 1.0;
        double est = 0.5 * (lo + hi);
        double frac = length(est) / total;
        while (fabs(frac - target) > 0.1 / double(S)) {
          if (frac > target) hi = est; else lo = est;
          est = 0.5 * (lo + hi);
          frac = length(est) / total;
        }
        su[n] = est;
      }
      sL = a.DL * (S - 1);
    }
    __syncthreads();
    m3 LR;
    for (int k = 0; k < 9; k++) LR.a[k] = sL * R.a[k];
    double cd[4], sd[3];
    cd[0] = 0.0;
    for (int n = 1; n < 4; n++) { cd[n] = -n * s[n - 1]; sd[n - 1] = n * c[n]; }
    for (int n = t; n < S; n += kShapeThreads) {
      if (n == 0) {
        st3(a.X, a.lds, g0, base);
        st3(a.Kd, a.lds, g0, d3{0.0, 0.0, 0.0});
        continue;
      }
      const double sv = su[n];
      const double w[3] = {sv, sv * sv, sv * sv * sv};
      const d3 pos = {fourier(bt.Ax, bt.Bx, c, s, w), fourier(bt.Ay, bt.By, c, s, w), 0.0};
      const d3 dir = {fourier(bt.Ax, bt.Bx, cd, sd, w), fourier(bt.Ay, bt.By, cd, sd, w), 0.0};
      st3(a.X, a.lds, g0 + n, base + mvec(LR, pos));
      st3(a.Kd, a.lds, g0 + n, mvec(LR, dir));
    }
  }
  __syncthreads();
  if (t == 0) {
    const d3 xb = ga3(b + CB_X);
    double kn = 0.0;
    d3 v1 = {0, 0, 0}, v2 = {0, 0, 0};
    for (int m = 0; m < S; m++) {
      const d3 k = ld3(a.Kd, a.lds, g0 + m);
      v1 = v1 + k;
      v2 = v2 + cross3(ld3(a.X, a.lds, g0 + m) - xb, k);
      kn += dot3(k, k);
    }
    double *fv = a.fvec + size_t(FV_REC) * f;
    fv[FV_KN] = kn;
    sa3(fv + FV_V1, v1);
    sa3(fv + FV_V2, v2);
  }
}

// Gauss-Jordan inverse with partial pivoting of a row-major 6x6 matrix
__device__ bool gj6(double *A) {
  int piv[6];
  for (int i = 0; i < 6; i++) piv[i] = i;
  for (int k = 0; k < 6; k++) {
    int p = k;
    double best = fabs(A[6 * k + k]);
    for (int i = k + 1; i < 6; i++)
      if (fabs(A[6 * i + k]) > best) { best = fabs(A[6 * i + k]); p = i; }
    if (best == 0.0) return false;
    if (p != k) {
      for (int j = 0; j < 6; j++) { const double x = A[6 * k + j]; A[6 * k + j] = A[6 * p + j]; A[6 * p + j] = x; }
      const int x = piv[k]; piv[k] = piv[p]; piv[p] = x;
    }
    const double inv = 1.0 / A[6 * k + k];
    A[6 * k + k] = 1.0;
    for (int j = 0; j < 6; j++) A[6 * k + j] *= inv;
    for (int i = 0; i < 6; i++) {
      if (i == k) continue;
      const double g = A[6 * i + k];
      if (g == 0.0) continue;
      A[6 * i + k] = 0.0;
      for (int j = 0; j < 6; j++) A[6 * i + j] -= g * A[6 * k + j];
    }
  }
  double B[36];
  for (int j = 0; j < 6; j++)
    for (int i = 0; i < 6; i++) B[6 * i + piv[j]] = A[6 * i + j];
  for (int k = 0; k < 36; k++) A[k] = B[k];
  return true;
}

// swimmer::make_precon_mat (swimmer.cpp:929-1027): (K^T M0^-1 K - C A^-1 B)^-1
// with M0 the self-mobility diagonal; the sums over particles use
// sum rcross(d) = rcross(sum d) and rcross(d)^2 = d d^T - |d|^2 I.
// One workgroup per swimmer; sets F_SING on a singular matrix.
__global__ __launch_bounds__(256) void k_cb_ktmk(CilDev a, const int *sw_seg0,
                                                 const int *sw_fil0, const int *sw_nfil,
                                                 const int *sw_blob0, const int *sw_nblob,
                                                 int *flags) {
  __shared__ double sh[4 * 21];
  const int n = blockIdx.x, t = threadIdx.x;
  double *b = a.body + size_t(n) * CB_REC;
  const d3 xb = ga3(b + CB_X);
  // r[0-2] sum d_s, r[3-8] sum d_s d_s^T (xx,xy,xz,yy,yz,zz), r[9] sum |d_s|^2,
  // r[10-19] the same for blobs, r[20-40] the filament Schur sums
  double r[41];
  double v[21];
  for (int k = 0; k < 10; k++) v[k] = 0.0;
  const int s0 = sw_seg0[n], ns = sw_nfil[n] * a.S;
  for (int i = t; i < ns; i += 256) {
    const d3 d = ld3(a.X, a.lds, s0 + i) - xb;
    v[0] += d.x; v[1] += d.y; v[2] += d.z;
    v[3] += d.x * d.x; v[4] += d.x * d.y; v[5] += d.x * d.z;
    v[6] += d.y * d.y; v[7] += d.y * d.z; v[8] += d.z * d.z;
    v[9] += dot3(d, d);
  }
  block_sums<10>(v, r, sh);
  for (int k = 0; k < 10; k++) v[k] = 0.0;
  const int b0 = sw_blob0[n], nb = sw_nblob[n];
  for (int m = t; m < nb; m += 256) {
    const d3 d = ld3(a.Db, a.ldb, b0 + m);
    v[0] += d.x; v[1] += d.y; v[2] += d.z;
    v[3] += d.x * d.x; v[4] += d.x * d.y; v[5] += d.x * d.z;
    v[6] += d.y * d.y; v[7] += d.y * d.z; v[8] += d.z * d.z;
    v[9] += dot3(d, d);
  }
  block_sums<10>(v, r + 10, sh);
  // 20-28 v1 v1^T, 29-37 v2 v1^T, 38-40 v2 v2^T diagonal; w: its off-diagonals
  double w2[3] = {0, 0, 0};
  if (a.dyn) {
    for (int k = 0; k < 21; k++) v[k] = 0.0;
    double w[3] = {0, 0, 0};
    const int f0 = sw_fil0[n], nf = sw_nfil[n];
    for (int i = t; i < nf; i += 256) {
      const double *fv = a.fvec + size_t(FV_REC) * (f0 + i);
      const double kn = fv[FV_KN];
      const double p[3] = {fv[FV_V1], fv[FV_V1 + 1], fv[FV_V1 + 2]};
      const double q[3] = {fv[FV_V2], fv[FV_V2 + 1], fv[FV_V2 + 2]};
      for (int rr = 0; rr < 3; rr++)
        for (int cc = 0; cc < 3; cc++) {
          v[3 * rr + cc] += p[rr] * p[cc] / kn;
          v[9 + 3 * rr + cc] += q[rr] * p[cc] / kn;
        }
      v[18] += q[0] * q[0] / kn;
      v[19] += q[1] * q[1] / kn;
      v[20] += q[2] * q[2] / kn;
      w[0] += q[0] * q[1] / kn;
      w[1] += q[0] * q[2] / kn;
      w[2] += q[1] * q[2] / kn;
    }
    block_sums<21>(v, r + 20, sh);
    block_sums<3>(w, w2, sh);
  }
  if (t != 0) return;
  const double sf = 6.0 * kPiC * a.mu * a.as, bf = 6.0 * kPiC * a.mu * a.ab;
  double M[36];
  for (int k = 0; k < 36; k++) M[k] = 0.0;
  const double eye = double(ns) * sf + double(nb) * bf;
  M[0] = M[7] = M[14] = eye;
  // upper-right: sf rcross(sum d_s) + bf rcross(sum d_b)
  const d3 ds = {r[0], r[1], r[2]}, db = {r[10], r[11], r[12]};
  const m3 rs = rcross(ds), rb = rcross(db);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      const double u = sf * rs.a[3 * i + j] + bf * rb.a[3 * i + j];
      M[6 * i + 3 + j] = u;
      M[6 * (3 + i) + j] = -u;
    }
  // lower-right: -sum rcross(d)^2 = sum (|d|^2 I - d d^T)
  const double Ss[9] = {r[3], r[4], r[5], r[4], r[6], r[7], r[5], r[7], r[8]};
  const double Sb[9] = {r[13], r[14], r[15], r[14], r[16], r[17], r[15], r[17], r[18]};
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      M[6 * (3 + i) + 3 + j] = sf * ((i == j ? r[9] : 0.0) - Ss[3 * i + j]) +
                               bf * ((i == j ? r[19] : 0.0) - Sb[3 * i + j]);
  if (a.dyn) {
    const double C22[9] = {r[38], w2[0], w2[1], w2[0], r[39], w2[2], w2[1], w2[2], r[40]};
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        M[6 * i + j] -= sf * r[20 + 3 * i + j];            // v1 v1^T
        M[6 * (3 + i) + j] -= sf * r[29 + 3 * i + j];      // v2 v1^T
        M[6 * i + 3 + j] -= sf * r[29 + 3 * j + i];        // its transpose
        M[6 * (3 + i) + 3 + j] -= sf * C22[3 * i + j];     // v2 v2^T
      }
  }
  if (!gj6(M)) flags[F_SING] = 1;
  for (int k = 0; k < 36; k++) b[CB_KI + k] = M[k];
}

// mobility_solver.cpp:223-306: cubic Hermite interpolation of the reference
// generalised force at `phase`, scaled by omega0 / 2 pi
__device__ double gen_force(const CilDev &a, double phase, double omega0) {
  const int N = a.ng;
  const double *g = a.gref;
  double pi = 0.5 * phase / kPiC;
  pi -= floor(pi);
  pi *= N;
  const int lo = int(pi);
  double q;
  if (lo == N) {
    q = g[0];
  } else {
    const int up = (lo == N - 1) ? 0 : lo + 1;
    const double ql = g[lo];
    const double qpl = (lo == 0) ? (g[1] - g[0]) : 0.5 * (g[up] - g[lo - 1]);
    const double qu = g[up];
    const double qpu = (up == 0) ? (g[0] - g[N - 1]) : 0.5 * (g[(up == N - 1) ? 0 : up + 1] - g[lo]);
    const double t = pi - lo, t2 = t * t, t3 = t2 * t;
    q = ql * (2.0 * t3 - 3.0 * t2 + 1.0) + qpl * (t3 - 2.0 * t2 + t) + qu * (3.0 * t2 - 2.0 * t3) +
        qpu * (t3 - t2);
  }
  return q * 0.5 * omega0 / kPiC;
}

// assemble_rhs (mobility_solver.cpp:196-359, prescribed branch), per tile
__global__ __launch_bounds__(256) void k_cil_rhs(CilDev a, const CTile *tiles, double *rhs) {
  const CTile T = tiles[blockIdx.x];
  const int t = threadIdx.x;
  const double *b = a.body + size_t(T.sw) * CB_REC;
  const d3 V = ga3(b + CB_VP), W = ga3(b + CB_VP + 3), xb = ga3(b + CB_X);
  if (t < T.n) {
    if (T.kind == 0) {
      const int i = T.m0 + t, f = T.f0 + t / a.S;
      d3 r = {0.0, 0.0, 0.0};
      if (!a.dyn) r = a.speed[f] * ld3(a.Kd, a.lds, i);
      if (!a.free_body) r = r + (V + cross3(W, ld3(a.X, a.lds, i) - xb));
      st3(rhs, a.lds, i, r);
    } else {
      const int m = T.m0 + t;
      d3 r = {0.0, 0.0, 0.0};
      if (!a.free_body) r = V + cross3(W, ld3(a.Db, a.ldb, m));
      st3(rhs + a.ob, a.ldb, m, r);
    }
  }
  if (T.kind == 0 && a.dyn && t < T.nf) {
    const int f = T.f0 + t;
    rhs[a.op + f] = -gen_force(a, a.phase[f], a.omega0[f]);
  }
}

// per-filament sums of the lanes' values in lds (a segment tile holds whole
// filaments; lane t < nf sums filament t's NSEG lanes in segment order)
__device__ double fil_sum(const double *sk, int t, int S) {
  double s = 0.0;
  for (int k = 0; k < S; k++) s += sk[t * S + k];
  return s;
}

// Preconditioner pass 1 (mobility_solver.cpp:427-560): o = M0^-1 in for the
// tile's particles, the tile's share of -K^T o, the filaments' phase values
// p' = (-in_phi - k.o) / |k|^2 (free bodies; fixed bodies: the final phase
// value (-in_phi - k.o) / (|k|^2 6 pi mu a)), and p' (v1, v2) added to the
// share.  `in` is the previous, unnormalised Arnoldi vector when
// norm_slot >= 0: its norm scal[norm_slot] is divided out here and the
// normalised vector written back (the tile owns its particles' rows and its
// filaments' phase rows, the swimmer's first tile its U rows).
__global__ __launch_bounds__(256) void k_cp1(CilDev a, const CTile *tiles, const int *sw_tile0,
                                             double *in, const double *scal, int norm_slot,
                                             double *part, double *pq, const int *flags) {
  if (flags && flags[F_DONE]) return;
  __shared__ double sk[256];
  __shared__ double sh[4 * 6];
  const CTile T = tiles[blockIdx.x];
  const int t = threadIdx.x;
  const double nrm = norm_slot >= 0 ? scal[norm_slot] : 1.0;
  const double sf = 6.0 * kPiC * a.mu * a.as, bf = 6.0 * kPiC * a.mu * a.ab;
  const double *b = a.body + size_t(T.sw) * CB_REC;
  const bool live = t < T.n;
  d3 o = {0, 0, 0}, d = {0, 0, 0};
  double kdot = 0.0;
  if (T.kind == 0) {
    const int i = T.m0 + t;
    if (live) {
      d3 v = ld3(in, a.lds, i);
      if (norm_slot >= 0) {
        v = {v.x / nrm, v.y / nrm, v.z / nrm};
        st3(in, a.lds, i, v);
      }
      o = sf * v;
      d = ld3(a.X, a.lds, i) - ga3(b + CB_X);
      if (a.dyn) kdot = dot3(ld3(a.Kd, a.lds, i), o);
    }
  } else {
    const int m = T.m0 + t;
    if (live) {
      d3 v = ld3(in + a.ob, a.ldb, m);
      if (norm_slot >= 0) {
        v = {v.x / nrm, v.y / nrm, v.z / nrm};
        st3(in + a.ob, a.ldb, m, v);
      }
      o = bf * v;
      d = ld3(a.Db, a.ldb, m);
    }
  }
  if (norm_slot >= 0 && a.free_body && blockIdx.x == sw_tile0[T.sw] && t < 6)
    in[a.oc + 6 * T.sw + t] /= nrm;
  double c[6] = {o.x, o.y, o.z, 0, 0, 0};
  const d3 tq = cross3(d, o);
  c[3] = tq.x; c[4] = tq.y; c[5] = tq.z;
  if (T.kind == 0 && a.dyn) {
    sk[t] = kdot;
    __syncthreads();
    if (t < T.nf) {
      const int f = T.f0 + t;
      const double tk = fil_sum(sk, t, a.S);
      double p = in[a.op + f];
      if (norm_slot >= 0) {
        p /= nrm;
        in[a.op + f] = p;
      }
      const double *fv = a.fvec + size_t(FV_REC) * f;
      if (a.free_body) {
        const double pp = (-p - tk) / fv[FV_KN];
        pq[f] = pp;
        for (int k = 0; k < 3; k++) { c[k] += pp * fv[FV_V1 + k]; c[3 + k] += pp * fv[FV_V2 + k]; }
      } else {
        pq[f] = (-p - tk) / (fv[FV_KN] * sf);
      }
    }
  }
  if (!a.free_body) return;
  double r[6];
  block_sums<6>(c, r, sh);
  if (t < 6) part[6 * blockIdx.x + t] = r[t];
}

// tile's share of [sum o; sum d x o] and, per filament, k . o: what the K
// products of the operator need from its input vector z
__device__ void tile_kt(const CilDev &a, const CTile &T, d3 o, d3 d, double kdot, double *zpart,
                        double *kz, double *sk, double *sh) {
  const int t = threadIdx.x;
  if (T.kind == 0 && a.dyn) {
    sk[t] = kdot;
    __syncthreads();
    if (t < T.nf) kz[T.f0 + t] = fil_sum(sk, t, a.S);
  }
  if (!a.free_body) return;
  const d3 tq = cross3(d, o);
  const double c[6] = {o.x, o.y, o.z, tq.x, tq.y, tq.z};
  double r[6];
  block_sums<6>(c, r, sh);
  if (t < 6) zpart[6 * blockIdx.x + t] = r[t];
}

// Preconditioner pass 2 (mobility_solver.cpp:560-650): each tile sums its
// swimmer's pass-1 shares (tile order), solves U = (K^T M0^-1 K)^-1 b,
// finishes its filaments' phase values p'' = p'/(6 pi mu a) - (v1.V +
// v2.W)/|k|^2 and writes out = M0^-1 (in + V + W x d + p'' k) for its
// particles.  Also the tile's shares for the K products of out.
__global__ __launch_bounds__(256) void k_cp2(CilDev a, const CTile *tiles, const int *sw_tile0,
                                             const int *sw_ntile, const double *part,
                                             const double *in, double *out, const double *pq,
                                             double *zpart, double *kz, const int *flags) {
  if (flags && flags[F_DONE]) return;
  __shared__ double sk[256];
  __shared__ double sh[4 * 6];
  __shared__ double sb[6], so[6];
  const CTile T = tiles[blockIdx.x];
  const int n = T.sw, t = threadIdx.x;
  const double sf = 6.0 * kPiC * a.mu * a.as, bf = 6.0 * kPiC * a.mu * a.ab;
  const double *b = a.body + size_t(n) * CB_REC;
  if (a.free_body) {
    if (t < 6) {
      double s = 0.0;
      for (int q = sw_tile0[n]; q < sw_tile0[n] + sw_ntile[n]; q++) s += part[6 * q + t];
      sb[t] = -in[a.oc + 6 * n + t] - s;
    }
    __syncthreads();
    if (t < 6) {
      double s = 0.0;
      for (int k = 0; k < 6; k++) s += b[CB_KI + 6 * t + k] * sb[k];
      so[t] = s;
      if (blockIdx.x == sw_tile0[n]) out[a.oc + 6 * n + t] = s;
    }
    __syncthreads();
  }
  const d3 V = a.free_body ? d3{so[0], so[1], so[2]} : d3{0, 0, 0};
  const d3 W = a.free_body ? d3{so[3], so[4], so[5]} : d3{0, 0, 0};
  const bool live = t < T.n;
  d3 o = {0, 0, 0}, d = {0, 0, 0};
  double kdot = 0.0;
  if (T.kind == 0) {
    const int i = T.m0 + t, f = T.f0 + t / a.S;
    double ph = 0.0;
    if (a.dyn && t < T.nf * a.S) {
      const double *fv = a.fvec + size_t(FV_REC) * f;
      ph = pq[f];
      if (a.free_body)
        ph = ph / sf - (dot3(ga3(fv + FV_V1), V) + dot3(ga3(fv + FV_V2), W)) / fv[FV_KN];
    }
    if (live) {
      d = ld3(a.X, a.lds, i) - ga3(b + CB_X);
      o = sf * ld3(in, a.lds, i);
      if (a.free_body) o = o + sf * (V + cross3(W, d));
      if (a.dyn) {
        const d3 k = ld3(a.Kd, a.lds, i);
        o = o + (sf * ph) * k;
      }
      st3(out, a.lds, i, o);
      if (a.dyn) kdot = dot3(ld3(a.Kd, a.lds, i), o);
    }
    if (a.dyn && t < T.nf) {
      const int g = T.f0 + t;
      double p2 = pq[g];
      if (a.free_body) {
        const double *fv = a.fvec + size_t(FV_REC) * g;
        p2 = p2 / sf - (dot3(ga3(fv + FV_V1), V) + dot3(ga3(fv + FV_V2), W)) / fv[FV_KN];
      }
      out[a.op + g] = p2;
    }
  } else {
    const int m = T.m0 + t;
    if (live) {
      d = ld3(a.Db, a.ldb, m);
      o = bf * ld3(in + a.ob, a.ldb, m);
      if (a.free_body) o = o + bf * (V + cross3(W, d));
      st3(out + a.ob, a.ldb, m, o);
    }
  }
  tile_kt(a, T, o, d, kdot, zpart, kz, sk, sh);
}

// The K shares of an operator input z that did not come from k_cp2.
__global__ __launch_bounds__(256) void k_ckt(CilDev a, const CTile *tiles, const double *z,
                                             double *zpart, double *kz, const int *flags) {
  if (flags && flags[F_DONE]) return;
  __shared__ double sk[256];
  __shared__ double sh[4 * 6];
  const CTile T = tiles[blockIdx.x];
  const int t = threadIdx.x;
  const double *b = a.body + size_t(T.sw) * CB_REC;
  d3 o = {0, 0, 0}, d = {0, 0, 0};
  double kdot = 0.0;
  if (t < T.n) {
    if (T.kind == 0) {
      const int i = T.m0 + t;
      o = ld3(z, a.lds, i);
      d = ld3(a.X, a.lds, i) - ga3(b + CB_X);
      if (a.dyn) kdot = dot3(ld3(a.Kd, a.lds, i), o);
    } else {
      const int m = T.m0 + t;
      o = ld3(z + a.ob, a.ldb, m);
      d = ld3(a.Db, a.ldb, m);
    }
  }
  tile_kt(a, T, o, d, kdot, zpart, kz, sk, sh);
}

// system_matrix_mult K products (mobility_solver.cpp:773-904): with the pair
// velocities vs = M_ss z_s + M_sb z_b and vb = M_bs z_s + M_bb z_b,
//   segment rows  vs - (V + W x d) - phi k
//   blob rows     vb - (V + W x d)
//   U rows        -sum [z; d x z]  (the tile shares, tile order)
//   phase rows    -k . z_s
__global__ __launch_bounds__(256) void k_ckprod(CilDev a, const CTile *tiles,
                                                const int *sw_tile0, const int *sw_ntile,
                                                const double *z, const double *vs,
                                                const double *vb, double *out,
                                                const double *zpart, const double *kz,
                                                const int *flags) {
  if (flags && flags[F_DONE]) return;
  const CTile T = tiles[blockIdx.x];
  const int n = T.sw, t = threadIdx.x;
  const double *b = a.body + size_t(n) * CB_REC;
  d3 V = {0, 0, 0}, W = {0, 0, 0};
  if (a.free_body) {
    V = ga3(z + a.oc + 6 * n);
    W = ga3(z + a.oc + 6 * n + 3);
    if (blockIdx.x == sw_tile0[n] && t < 6) {
      double s = 0.0;
      for (int q = sw_tile0[n]; q < sw_tile0[n] + sw_ntile[n]; q++) s += zpart[6 * q + t];
      out[a.oc + 6 * n + t] = -s;
    }
  }
  if (T.kind == 0 && a.dyn && t < T.nf) out[a.op + T.f0 + t] = -kz[T.f0 + t];
  if (t >= T.n) return;
  if (T.kind == 0) {
    const int i = T.m0 + t;
    d3 v = ld3(vs, a.lds, i);
    if (a.free_body) v = v - (V + cross3(W, ld3(a.X, a.lds, i) - ga3(b + CB_X)));
    if (a.dyn) v = v - z[a.op + T.f0 + t / a.S] * ld3(a.Kd, a.lds, i);
    st3(out, a.lds, i, v);
  } else {
    const int m = T.m0 + t;
    d3 v = ld3(vb, a.ldb, m);
    if (a.free_body) v = v - (V + cross3(W, ld3(a.Db, a.ldb, m)));
    st3(out + a.ob, a.ldb, m, v);
  }
}

// cilia_sim_main.cpp:218-232: x += dt V, u = dt W, q = exp(u) q_{n-1}
__global__ void k_cb_integrate(CilDev a, const double *x) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= a.Nsw) return;
  double *b = a.body + size_t(n) * CB_REC;
  const double *U = a.free_body ? x + a.oc + 6 * n : b + CB_VP;
  for (int k = 0; k < 3; k++) b[CB_X + k] += a.dt * U[k];
  const d3 u = {a.dt * U[3], a.dt * U[4], a.dt * U[5]};
  sa3(b + CB_U, u);
  sa4(b + CB_Q, qmul(lie_exp(u), ga4(b + CB_QM1)));
}

// the solved phase speeds (mobility_solver.cpp:1306-1322)
__global__ void k_cf_store_speed(CilDev a, const double *x) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= a.Nf || !a.dyn) return;
  a.speed[f] = x[a.op + f];
}

// filament::end_of_step (filament.cpp:1020-1031): phase += speed dt, frame
// q_{n-1} = q
__global__ void k_cf_end(CilDev a) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= a.Nf) return;
  a.phase[f] += a.speed[f] * a.dt;
  for (int k = 0; k < 4; k++) a.fqm1[4 * f + k] = a.fq[4 * f + k];
}

// warm start of the phase speeds: omega0 (filament.cpp:829)
__global__ void k_cf_speed_guess(CilDev a, double *x) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= a.Nf || !a.dyn) return;
  x[a.op + f] = a.omega0[f];
}

// vel_dir . f per filament, in the reference's order (cilia_sim_main.cpp:
// 234-246)
__global__ void k_cf_genforce(CilDev a, const double *x, double *out) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= a.Nf) return;
  double g = 0.0;
  for (int m = 0; m < a.S; m++) {
    const int i = f * a.S + m;
    g += a.Kd[i] * x[i];
    g += a.Kd[a.lds + i] * x[a.lds + i];
    g += a.Kd[2 * a.lds + i] * x[2 * a.lds + i];
  }
  out[f] = g;
}

// reference layout <-> device layout of a system vector
__global__ void k_cil_from_ref(CilDev a, const double *ref, double *dev) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int ns3 = 3 * a.Ns, nb3 = 3 * a.Nb;
  const int nu = a.free_body ? 6 * a.Nsw : 0, np = a.dyn ? a.Nf : 0;
  if (e < ns3) dev[(e % 3) * a.lds + e / 3] = ref[e];
  else if (e < ns3 + nb3) dev[a.ob + ((e - ns3) % 3) * a.ldb + (e - ns3) / 3] = ref[e];
  else if (e < ns3 + nb3 + nu) dev[a.oc + (e - ns3 - nb3)] = ref[e];
  else if (e < ns3 + nb3 + nu + np) dev[a.op + (e - ns3 - nb3 - nu)] = ref[e];
}
__global__ void k_cil_to_ref(CilDev a, const double *dev, double *ref) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int ns3 = 3 * a.Ns, nb3 = 3 * a.Nb;
  const int nu = a.free_body ? 6 * a.Nsw : 0, np = a.dyn ? a.Nf : 0;
  if (e < ns3) ref[e] = dev[(e % 3) * a.lds + e / 3];
  else if (e < ns3 + nb3) ref[e] = dev[a.ob + ((e - ns3) % 3) * a.ldb + (e - ns3) / 3];
  else if (e < ns3 + nb3 + nu) ref[e] = dev[a.oc + (e - ns3 - nb3)];
  else if (e < ns3 + nb3 + nu + np) ref[e] = dev[a.op + (e - ns3 - nb3 - nu)];
}

// ---- host helpers ------------------------------------------------------------

struct HQ { double s, x, y, z; };
HQ hq_mul(const HQ &p, const HQ &q) {
  return {p.s * q.s - (p.x * q.x + p.y * q.y + p.z * q.z), p.s * q.x + q.s * p.x + p.y * q.z - p.z * q.y,
          p.s * q.y + q.s * p.y + p.z * q.x - p.x * q.z, p.s * q.z + q.s * p.z + p.x * q.y - p.y * q.x};
}
HQ hq_sqrt(HQ q) {
  if (q.s < -0.999999999) return {0.0, 0.0, 0.0, 1.0};
  const double s = std::sqrt(0.5 * (1.0 + q.s)), t = 2.0 * s;
  return {s, q.x / t, q.y / t, q.z / t};
}

}  // namespace

// ==========================================================================

struct CiliaHost : KrylovState {
  CilDev d;
  int precision = GMI_FP64;
  double tol = 1e-10;
  int maxit = 150;
  int nsys_ref = 0;
  std::vector<void *> owned;
  CTile *tiles = nullptr;
  int ntiles = 0;
  int *sw_tile0 = nullptr, *sw_ntile = nullptr;
  int *sw_seg0 = nullptr, *sw_fil0 = nullptr, *sw_nfil = nullptr, *sw_blob0 = nullptr,
      *sw_nblob = nullptr;
  double *tpart = nullptr, *zpart = nullptr, *pq = nullptr, *kz = nullptr, *gout = nullptr,
         *refbuf = nullptr;
  int last_iters = 0;
};

static CiliaHost *cil_of(gmi_ctx *c) {
  return (c && c->extra && c->extra_kind == kExtraCilia) ? static_cast<CiliaHost *>(c->extra)
                                                               : nullptr;
}

static void cilia_release(gmi_ctx *c) {
  CiliaHost *s = static_cast<CiliaHost *>(c->extra);
  if (!s) return;
  s->drop_graphs();
  for (void *p : s->owned) (void)hipFree(p);
  if (s->h_flags) (void)hipHostFree(s->h_flags);
  if (s->h_scal) (void)hipHostFree(s->h_scal);
  c->tt_only = false;
  delete s;
}

template <typename T>
static int calloc_dev(gmi_ctx *c, CiliaHost *s, T **p, size_t n) {
  void *q = nullptr;
  int rc = hip_check(c, hipMalloc(&q, sizeof(T) * std::max<size_t>(n, 1)), "hipMalloc cilia");
  if (rc) return rc;
  s->owned.push_back(q);
  *p = static_cast<T *>(q);
  return hip_check(c, hipMemsetAsync(q, 0, sizeof(T) * std::max<size_t>(n, 1), c->stream),
                   "memset cilia");
}

template <typename T>
static int put(gmi_ctx *c, T *dst, const T *src, size_t n) {
  if (n == 0) return GMI_OK;
  return hip_check(c, hipMemcpyAsync(dst, src, sizeof(T) * n, hipMemcpyHostToDevice, c->stream),
                   "upload cilia");
}

// The system as the Krylov engine's operator.
struct CiliaOp {
  gmi_ctx *c;
  CiliaHost *s;
  int mult(double *z, double *out, int mode, const int *done, bool zpart_ready) {
    CilDev &d = s->d;
    if (!zpart_ready)
      hipLaunchKernelGGL(k_ckt, dim3(s->ntiles), dim3(256), 0, c->stream, d, s->tiles, z,
                         s->zpart, s->kz, done);
    const WorkList &wl = c->lists[kSS | kSB | kBS | kBB];
    PairArgs pa;
    pa.Xs = c->Xs; pa.Fs = z; pa.lds = c->lds;
    pa.Xb = c->Xb; pa.Fb = z + d.ob; pa.ldb = c->ldb;
    pa.Vs = c->Vs; pa.Vb = c->Vb;
    pa.part = c->part; pa.tiles = wl.d_tiles; pa.work = wl.d_work;
    pa.mask = kSS | kSB | kBS | kBB; pa.accumulate = 0;
    pa.a_s = c->as; pa.a_b = c->ab; pa.mu = c->mu;
    pa.tt_only = 1;
    const int nred = wl.nsplit_tiles ? wl.ntiles : 0;
    int rc;
    if (c->wall) {
      // FP64 only (the stepper refuses the mixed modes above the wall)
      pa.skip = done;
      if ((rc = hip_check(c, launch_pair_wall(pa, wl.nwork, nred, c->stream), "cilia wall pairs")))
        return rc;
    } else if (mode == 2) {
      // one launch: FP32 once the relaxed switch cleared F_SKIP32, else FP64
      pa.skip = s->flags + F_DONE;
      pa.sel32 = s->flags + F_SKIP32;
      if ((rc = hip_check(c, launch_pair_dyn(pa, wl.nwork, nred, c->stream), "cilia pairs")))
        return rc;
    } else {
      pa.skip = done;
      if ((rc = hip_check(c, launch_pair(pa, wl.nwork, nred, mode == 1, c->stream), "cilia pairs")))
        return rc;
    }
    hipLaunchKernelGGL(k_ckprod, dim3(s->ntiles), dim3(256), 0, c->stream, d, s->tiles,
                       s->sw_tile0, s->sw_ntile, z, c->Vs, c->Vb, out, s->zpart, s->kz, done);
    return launch_ok(c, "cilia K products");
  }
  bool fp64_krylov() const { return false; }
  bool zero_start() const { return false; }
  // Arnoldi steps go through the engine's generic kernels
  bool fused() const { return false; }
  int fused_parts() const { return 0; }
  int fused_step(KrylovState *, int, bool, bool, const GivensArgs &) { return GMI_ERR_STATE; }
  int precond(double *in, double *out, const int *skip, int norm_slot) {
    CilDev &d = s->d;
    hipLaunchKernelGGL(k_cp1, dim3(s->ntiles), dim3(256), 0, c->stream, d, s->tiles, s->sw_tile0,
                       in, s->scal, norm_slot, s->tpart, s->pq, skip);
    hipLaunchKernelGGL(k_cp2, dim3(s->ntiles), dim3(256), 0, c->stream, d, s->tiles, s->sw_tile0,
                       s->sw_ntile, s->tpart, in, out, s->pq, s->zpart, s->kz, skip);
    return launch_ok(c, "cilia preconditioner");
  }
};

static int cil_sync(gmi_ctx *c) { return hip_check(c, hipStreamSynchronize(c->stream), "sync"); }

// device -> host on the context's (non-blocking) stream, then wait
static int d2h(gmi_ctx *c, void *dst, const void *src, size_t bytes) {
  int rc;
  if ((rc = hip_check(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream), "D2H")))
    return rc;
  return cil_sync(c);
}

static int initial_guess(gmi_ctx *c, CiliaHost *s) {
  CilDev &d = s->d;
  hipLaunchKernelGGL(k_cb_guess, dim3(nblk(d.Nsw, 64)), dim3(64), 0, c->stream, d);
  if (d.Nb > 0)
    hipLaunchKernelGGL(k_cb_blobs, dim3(nblk(d.Nb, 256)), dim3(256), 0, c->stream, d);
  hipLaunchKernelGGL(k_cf_shape, dim3(d.Nf), dim3(kShapeThreads), 0, c->stream, d);
  hipLaunchKernelGGL(k_cf_speed_guess, dim3(nblk(d.Nf, 64)), dim3(64), 0, c->stream, d, s->x);
  int rc;
  if ((rc = launch_ok(c, "cilia shapes"))) return rc;
  if (!d.free_body) return GMI_OK;
  hipLaunchKernelGGL(k_set_ints, dim3(1), dim3(64), 0, c->stream, s->flags + F_SING, 1, 0,
                     nullptr);
  hipLaunchKernelGGL(k_cb_ktmk, dim3(d.Nsw), dim3(256), 0, c->stream, d, s->sw_seg0, s->sw_fil0,
                     s->sw_nfil, s->sw_blob0, s->sw_nblob, s->flags);
  if ((rc = launch_ok(c, "cilia preconditioner matrices"))) return rc;
  if ((rc = krylov_read_flags(c, s))) return rc;
  if (s->h_flags[F_SING]) return fail(c, GMI_ERR_SINGULAR, "singular K^T M0^-1 K");
  return GMI_OK;
}

static int compute_velocities(gmi_ctx *c, CiliaHost *s, int *its) {
  CilDev &d = s->d;
  hipLaunchKernelGGL(k_cil_rhs, dim3(s->ntiles), dim3(256), 0, c->stream, d, s->tiles, s->rhs);
  int rc;
  if ((rc = launch_ok(c, "cilia rhs"))) return rc;
  CiliaOp op{c, s};
  int it = 0;
  if ((rc = krylov_solve(c, s, op, d.n, std::min(s->nsys_ref, s->maxit), s->precision, s->tol, &it)))
    return rc;
  hipLaunchKernelGGL(k_cf_store_speed, dim3(nblk(d.Nf, 64)), dim3(64), 0, c->stream, d, s->x);
  if ((rc = launch_ok(c, "phase speeds"))) return rc;
  s->last_iters = it;
  if (its) *its = it;
  return GMI_OK;
}

// host reference-layout vector -> device layout in dst (uses s->refbuf)
static int ref_to_dev(gmi_ctx *c, CiliaHost *s, const double *ref, double *dst) {
  int rc;
  if ((rc = hip_check(c, hipMemsetAsync(dst, 0, sizeof(double) * s->d.n, c->stream), "zero")))
    return rc;
  if ((rc = put(c, s->refbuf, ref, size_t(s->nsys_ref)))) return rc;
  hipLaunchKernelGGL(k_cil_from_ref, dim3(nblk(s->nsys_ref, 256)), dim3(256), 0, c->stream, s->d,
                     s->refbuf, dst);
  return launch_ok(c, "to device layout");
}

static int dev_to_ref(gmi_ctx *c, CiliaHost *s, const double *src, double *ref) {
  hipLaunchKernelGGL(k_cil_to_ref, dim3(nblk(s->nsys_ref, 256)), dim3(256), 0, c->stream, s->d, src,
                     s->refbuf);
  int rc;
  if ((rc = launch_ok(c, "to reference layout"))) return rc;
  if ((rc = hip_check(c, hipMemcpyAsync(ref, s->refbuf, sizeof(double) * s->nsys_ref,
                                        hipMemcpyDeviceToHost, c->stream), "D2H")))
    return rc;
  return cil_sync(c);
}

static int soa_to_host(gmi_ctx *c, const double *src, int ld, int n, double *dst) {
  std::vector<double> h(3 * size_t(ld));
  int rc;
  if ((rc = hip_check(c, hipMemcpyAsync(h.data(), src, sizeof(double) * 3 * ld,
                                        hipMemcpyDeviceToHost, c->stream), "D2H")))
    return rc;
  if ((rc = cil_sync(c))) return rc;
  for (int i = 0; i < n; i++)
    for (int k = 0; k < 3; k++) dst[3 * size_t(i) + k] = h[k * size_t(ld) + i];
  return GMI_OK;
}

}  // namespace gmi

using namespace gmi;

extern "C" {

int gmi_cilia_create(gmi_ctx *c, const gmi_cilia_params *p) {
  if (!c || !p || !p->nfil || !p->fil_refs) return GMI_ERR_ARG;
  if (c->extra) return fail(c, GMI_ERR_STATE, "context already has a time stepper");
  const bool wall = c->wall;
  if (wall && (c->nsw != 1 || !p->fixed_body || p->body_velocity ||
               (p->body_x && (p->body_x[0] != 0.0 || p->body_x[1] != 0.0 || p->body_x[2] != 0.0))))
    return fail(c, GMI_ERR_ARG,
                "a carpet above the wall is one surface at rest at the origin: one group, "
                "fixed_body, no body velocity");
  if (wall && c->precision != GMI_FP64)
    return fail(c, GMI_ERR_ARG, "the wall-bounded mobility is FP64 only");
  const int S = c->nseg_fil;
  if (S < 2 || S > kMaxSegPerFil) return fail(c, GMI_ERR_ARG, "NSEG must be in [2, 256]");
  int Nf = 0;
  for (int n = 0; n < c->nsw; n++) {
    if (p->nfil[n] < 1 || p->nfil[n] * S != c->nseg[n])
      return fail(c, GMI_ERR_ARG, "nfil*NSEG must match the context's segment counts");
    Nf += p->nfil[n];
  }
  if (c->Nb > 0 && !p->blob_refs) return fail(c, GMI_ERR_ARG, "blob_refs required");
  if (!(p->dt > 0) || !(p->gmres_tol > 0) || p->gmres_max_iter < 1 ||
      p->gmres_max_iter > kMaxCols - 1)
    return fail(c, GMI_ERR_ARG, "bad time step or GMRES parameters");
  if (p->beat < GMI_BEAT_BUILD || p->beat > GMI_BEAT_CORAL_LARVAE)
    return fail(c, GMI_ERR_ARG, "unknown beat");
  if (p->beat == GMI_BEAT_BUILD &&
      (!(p->amplitude > 0 && p->amplitude <= 2) || !(p->recovery_window > 0) ||
       !(p->effective_stroke > 0) || !(p->recovery_window + p->effective_stroke < 1)))
    return fail(c, GMI_ERR_ARG, "bad build-a-beat parameters");
  if (p->dynamic_phase && (!p->gen_force_refs || p->n_gen_force_refs < 2))
    return fail(c, GMI_ERR_ARG, "dynamic phases need at least two reference generalised forces");
  if (!wall && !p->phase && !(p->r_over_l > 0))
    return fail(c, GMI_ERR_ARG, "R_OVER_L must be positive");
  int rc;
  if ((rc = hip_check(c, hipSetDevice(c->device), "set device"))) return rc;
  CiliaHost *s = new CiliaHost();
  CilDev &d = s->d;
  d.S = S; d.Nf = Nf; d.Nsw = c->nsw; d.Ns = c->Ns; d.Nb = c->Nb; d.lds = c->lds; d.ldb = c->ldb;
  d.free_body = p->fixed_body ? 0 : 1;
  d.dyn = p->dynamic_phase ? 1 : 0;
  d.ob = 3 * d.lds;
  d.oc = d.ob + 3 * d.ldb;
  d.op = d.oc + (d.free_body ? 6 * d.Nsw : 0);
  d.n = d.op + (d.dyn ? ((Nf + 63) / 64) * 64 : 0);
  d.n = ((d.n + 63) / 64) * 64;
  d.mu = c->mu; d.as = c->as; d.ab = c->ab; d.dt = p->dt; d.DL = 2.2 * c->as;
  d.X = c->Xs; d.Xb = c->Xb;
  s->nsys_ref = 3 * d.Ns + 3 * d.Nb + (d.free_body ? 6 * d.Nsw : 0) + (d.dyn ? Nf : 0);
  s->tol = p->gmres_tol;
  s->maxit = p->gmres_max_iter;
  s->precision = c->precision;
  Beat &bt = d.beat;
  std::memset(&bt, 0, sizeof bt);
  bt.kind = p->beat;
  if (p->beat == GMI_BEAT_BUILD) {
    bt.theta = std::acos(0.5 * p->amplitude);
    bt.wR = p->recovery_window * 2.0 * kPiC;
    bt.wE = p->effective_stroke * 2.0 * kPiC;
    bt.zva = p->zero_velocity_avoidance;
  } else {
    const bool fb = p->beat == GMI_BEAT_FULFORD_BLAKE;
    std::memcpy(bt.Ax, fb ? kFbAx : kClAx, sizeof bt.Ax);
    std::memcpy(bt.Ay, fb ? kFbAy : kClAy, sizeof bt.Ay);
    std::memcpy(bt.Bx, fb ? kFbBx : kClBx, sizeof bt.Bx);
    std::memcpy(bt.By, fb ? kFbBy : kClBy, sizeof bt.By);
  }
  auto bail = [&](int code) {
    c->extra = s;
    cilia_release(c);
    c->extra = nullptr;
    return code;
  };
  // ---- host-side set-up (swimmer.cpp:160-237, filament.cpp:197-290) ----
  std::vector<int> fil_sw(Nf), blob_sw(std::max(d.Nb, 1)), seg0(d.Nsw), fil0(d.Nsw),
      nfv(d.Nsw), blob0(d.Nsw), nblv(d.Nsw);
  std::vector<double> body(size_t(d.Nsw) * CB_REC, 0.0), fref(3 * size_t(Nf)), phase(Nf),
      omega0(Nf), fq(4 * size_t(Nf)), bref(3 * size_t(d.ldb), 0.0);
  int fo = 0;
  for (int n = 0; n < d.Nsw; n++) {
    seg0[n] = c->seg_off[n]; fil0[n] = fo; nfv[n] = p->nfil[n];
    blob0[n] = c->blob_off[n]; nblv[n] = c->nblob[n];
    double *b = &body[size_t(n) * CB_REC];
    for (int k = 0; k < 3; k++) b[CB_X + k] = p->body_x ? p->body_x[3 * n + k] : 0.0;
    b[CB_Q] = b[CB_QM1] = 1.0;
    if (p->fixed_body && p->body_velocity)
      for (int k = 0; k < 6; k++) b[CB_VP + k] = p->body_velocity[6 * n + k];
    for (int m = 0; m < c->nblob[n]; m++) {
      const int g = c->blob_off[n] + m;
      blob_sw[g] = n;
      for (int k = 0; k < 3; k++) bref[k * size_t(d.ldb) + g] = p->blob_refs[3 * size_t(g) + k];
    }
    for (int i = 0; i < p->nfil[n]; i++, fo++) {
      fil_sw[fo] = n;
      const double *r = p->fil_refs + 3 * size_t(fo);
      for (int k = 0; k < 3; k++) fref[3 * size_t(fo) + k] = r[k];
      const double rn = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
      if (!wall && !(rn > 0)) {
        delete s;
        return fail(c, GMI_ERR_ARG, "filament base at the body centre");
      }
      if (wall && !(r[2] > 0)) {
        delete s;
        return fail(c, GMI_ERR_ARG, "filament base not above the wall");
      }
      // beat frame: x to the surface normal -- the base direction on a body,
      // +z above the wall (swimmer.cpp:91) -- and the second frame vector to
      // the polar direction (filament.cpp:245-282)
      const double dir[3] = {wall ? 0.0 : r[0] / rn, wall ? 0.0 : r[1] / rn,
                             wall ? 1.0 : r[2] / rn};
      const HQ qt = hq_sqrt(HQ{dir[0], 0.0, -dir[2], dir[1]});
      const double bn[3] = {2.0 * (qt.x * qt.y - qt.s * qt.z), 1.0 - 2.0 * (qt.x * qt.x + qt.z * qt.z),
                            2.0 * (qt.y * qt.z + qt.s * qt.x)};
      double pd[3] = {0.0, 0.0, -1.0};
      const double dp = -dir[2];
      for (int k = 0; k < 3; k++) pd[k] -= dp * dir[k];
      const double pn = std::sqrt(pd[0] * pd[0] + pd[1] * pd[1] + pd[2] * pd[2]);
      HQ q2{1.0, 0.0, 0.0, 0.0};
      if (pn > 0.0) {
        for (int k = 0; k < 3; k++) pd[k] /= pn;
        q2 = hq_sqrt(HQ{pd[0] * bn[0] + pd[1] * bn[1] + pd[2] * bn[2], bn[1] * pd[2] - bn[2] * pd[1],
                        bn[2] * pd[0] - bn[0] * pd[2], bn[0] * pd[1] - bn[1] * pd[0]});
      }
      const HQ q = hq_mul(q2, qt);
      fq[4 * size_t(fo)] = q.s; fq[4 * size_t(fo) + 1] = q.x;
      fq[4 * size_t(fo) + 2] = q.y; fq[4 * size_t(fo) + 3] = q.z;
      if (p->phase) {
        phase[fo] = p->phase[fo];
      } else if (wall) {
        // filament.cpp:222: a wave of one filament length along x
        phase[fo] = 2.0 * kPiC * r[0] / (d.DL * (S - 1)) + 2.0 * kPiC * 116.0 / 300.0;
      } else {
        const double phi = std::atan2(b[CB_X + 1] + r[1], b[CB_X] + r[0]);
        double ph = phi * (2.0 * kPiC * p->r_over_l);
        ph -= 2.0 * kPiC * std::floor(0.5 * ph / kPiC);
        phase[fo] = ph;
      }
      omega0[fo] = p->omega0 ? p->omega0[fo] : 2.0 * kPiC;
    }
  }
  // tiles: per swimmer, whole filaments then blobs
  std::vector<CTile> tiles;
  std::vector<int> t0(d.Nsw), tn(d.Nsw);
  const int fpt = std::max(1, 256 / S);
  for (int n = 0; n < d.Nsw; n++) {
    t0[n] = int(tiles.size());
    for (int i = 0; i < p->nfil[n]; i += fpt) {
      CTile T;
      T.sw = n; T.kind = 0;
      T.f0 = fil0[n] + i;
      T.nf = std::min(fpt, p->nfil[n] - i);
      T.m0 = T.f0 * S;
      T.n = T.nf * S;
      tiles.push_back(T);
    }
    for (int m = 0; m < c->nblob[n]; m += 256) {
      CTile T;
      T.sw = n; T.kind = 1;
      T.m0 = c->blob_off[n] + m;
      T.n = std::min(256, c->nblob[n] - m);
      T.f0 = 0; T.nf = 0;
      tiles.push_back(T);
    }
    tn[n] = int(tiles.size()) - t0[n];
  }
  s->ntiles = int(tiles.size());
  // ---- device allocations ----
  const int cols = p->gmres_max_iter;
  s->cols = cols;
  const size_t np = size_t(d.n);
  int *dfil_sw, *dblob_sw;
  double *dfref, *dbref, *dgref = nullptr;
  if ((rc = calloc_dev(c, s, &dfil_sw, Nf)) || (rc = calloc_dev(c, s, &dblob_sw, blob_sw.size())) ||
      (rc = calloc_dev(c, s, &dfref, fref.size())) || (rc = calloc_dev(c, s, &dbref, bref.size())) ||
      (rc = calloc_dev(c, s, &d.Db, 3 * size_t(d.ldb))) ||
      (rc = calloc_dev(c, s, &d.Kd, 3 * size_t(d.lds))) ||
      (rc = calloc_dev(c, s, &d.phase, Nf)) || (rc = calloc_dev(c, s, &d.omega0, Nf)) ||
      (rc = calloc_dev(c, s, &d.speed, Nf)) || (rc = calloc_dev(c, s, &d.fq, 4 * size_t(Nf))) ||
      (rc = calloc_dev(c, s, &d.fqm1, 4 * size_t(Nf))) ||
      (rc = calloc_dev(c, s, &d.fvec, size_t(FV_REC) * Nf)) ||
      (rc = calloc_dev(c, s, &d.body, body.size())) ||
      (rc = calloc_dev(c, s, &s->tiles, tiles.size())) ||
      (rc = calloc_dev(c, s, &s->sw_tile0, d.Nsw)) || (rc = calloc_dev(c, s, &s->sw_ntile, d.Nsw)) ||
      (rc = calloc_dev(c, s, &s->sw_seg0, d.Nsw)) || (rc = calloc_dev(c, s, &s->sw_fil0, d.Nsw)) ||
      (rc = calloc_dev(c, s, &s->sw_nfil, d.Nsw)) || (rc = calloc_dev(c, s, &s->sw_blob0, d.Nsw)) ||
      (rc = calloc_dev(c, s, &s->sw_nblob, d.Nsw)) ||
      (rc = calloc_dev(c, s, &s->tpart, 6 * size_t(s->ntiles))) ||
      (rc = calloc_dev(c, s, &s->zpart, 6 * size_t(s->ntiles))) ||
      (rc = calloc_dev(c, s, &s->pq, Nf)) || (rc = calloc_dev(c, s, &s->kz, Nf)) ||
      (rc = calloc_dev(c, s, &s->gout, Nf)) ||
      (rc = calloc_dev(c, s, &s->refbuf, size_t(s->nsys_ref))) ||
      (rc = calloc_dev(c, s, &s->x, np)) || (rc = calloc_dev(c, s, &s->rhs, np)) ||
      (rc = calloc_dev(c, s, &s->Qk, np * (cols + 1))) ||
      (rc = calloc_dev(c, s, &s->H, size_t(cols + 1) * (cols + 1))) ||
      (rc = calloc_dev(c, s, &s->cs, cols + 1)) || (rc = calloc_dev(c, s, &s->sn, cols + 1)) ||
      (rc = calloc_dev(c, s, &s->beta, cols + 1)) ||
      (rc = calloc_dev(c, s, &s->bpre, cols + 1)) || (rc = calloc_dev(c, s, &s->h1, cols + 1)) ||
      (rc = calloc_dev(c, s, &s->h2, cols + 1)) || (rc = calloc_dev(c, s, &s->y, cols + 1)) ||
      (rc = calloc_dev(c, s, &s->z, np)) || (rc = calloc_dev(c, s, &s->tmp, np)) ||
      (rc = calloc_dev(c, s, &s->scal, 16)) || (rc = calloc_dev(c, s, &s->part, nblk(d.n, 256))) ||
      (rc = calloc_dev(c, s, &s->flags, F_N)))
    return bail(rc);
  s->pp_half = size_t(cols + 1) * dot_shape(d.n).nb;
  if ((rc = calloc_dev(c, s, &s->pp, 2 * s->pp_half))) return bail(rc);
  if (d.dyn) {
    if ((rc = calloc_dev(c, s, &dgref, size_t(p->n_gen_force_refs)))) return bail(rc);
    if ((rc = put(c, dgref, p->gen_force_refs, size_t(p->n_gen_force_refs)))) return bail(rc);
    d.ng = p->n_gen_force_refs;
  }
  d.gref = dgref;
  if ((rc = hip_check(c, hipHostMalloc((void **)&s->h_flags, sizeof(int) * F_N,
                                       hipHostMallocDefault), "pinned flags")) ||
      (rc = hip_check(c, hipHostMalloc((void **)&s->h_scal, sizeof(double) * 16,
                                       hipHostMallocDefault), "pinned scal")))
    return bail(rc);
  d.fil_sw = dfil_sw; d.blob_sw = dblob_sw; d.fref = dfref; d.bref = dbref;
  if ((rc = put(c, dfil_sw, fil_sw.data(), fil_sw.size())) ||
      (rc = put(c, dblob_sw, blob_sw.data(), blob_sw.size())) ||
      (rc = put(c, dfref, fref.data(), fref.size())) || (rc = put(c, dbref, bref.data(), bref.size())) ||
      (rc = put(c, d.phase, phase.data(), phase.size())) ||
      (rc = put(c, d.omega0, omega0.data(), omega0.size())) ||
      (rc = put(c, d.speed, omega0.data(), omega0.size())) ||
      (rc = put(c, d.fq, fq.data(), fq.size())) || (rc = put(c, d.fqm1, fq.data(), fq.size())) ||
      (rc = put(c, d.body, body.data(), body.size())) ||
      (rc = put(c, s->tiles, tiles.data(), tiles.size())) ||
      (rc = put(c, s->sw_tile0, t0.data(), t0.size())) || (rc = put(c, s->sw_ntile, tn.data(), tn.size())) ||
      (rc = put(c, s->sw_seg0, seg0.data(), seg0.size())) ||
      (rc = put(c, s->sw_fil0, fil0.data(), fil0.size())) ||
      (rc = put(c, s->sw_nfil, nfv.data(), nfv.size())) ||
      (rc = put(c, s->sw_blob0, blob0.data(), blob0.size())) ||
      (rc = put(c, s->sw_nblob, nblv.data(), nblv.size())) || (rc = cil_sync(c)))
    return bail(rc);
  c->extra = s;
  c->extra_release = cilia_release;
  c->extra_kind = kExtraCilia;
  c->tt_only = true;
  return GMI_OK;
}

int gmi_cilia_initial_guess(gmi_ctx *c, int nt) {
  (void)nt;
  CiliaHost *s = cil_of(c);
  if (!s) return c ? fail(c, GMI_ERR_STATE, "no prescribed-cilia stepper") : GMI_ERR_ARG;
  return initial_guess(c, s);
}

int gmi_cilia_compute_velocities(gmi_ctx *c, int nt, int *its) {
  (void)nt;
  CiliaHost *s = cil_of(c);
  if (!s) return c ? fail(c, GMI_ERR_STATE, "no prescribed-cilia stepper") : GMI_ERR_ARG;
  return compute_velocities(c, s, its);
}

int gmi_cilia_integrate_bodies(gmi_ctx *c) {
  CiliaHost *s = cil_of(c);
  if (!s) return c ? fail(c, GMI_ERR_STATE, "no prescribed-cilia stepper") : GMI_ERR_ARG;
  hipLaunchKernelGGL(k_cb_integrate, dim3(nblk(s->d.Nsw, 64)), dim3(64), 0, c->stream, s->d, s->x);
  return launch_ok(c, "cilia body update");
}

int gmi_cilia_end_of_step(gmi_ctx *c, int nt) {
  (void)nt;
  CiliaHost *s = cil_of(c);
  if (!s) return c ? fail(c, GMI_ERR_STATE, "no prescribed-cilia stepper") : GMI_ERR_ARG;
  hipLaunchKernelGGL(k_cf_end, dim3(nblk(s->d.Nf, 64)), dim3(64), 0, c->stream, s->d);
  int rc;
  if ((rc = launch_ok(c, "cilia end of step"))) return rc;
  return cil_sync(c);
}

int gmi_cilia_step(gmi_ctx *c, int nt, int *its) {
  int rc;
  if ((rc = gmi_cilia_initial_guess(c, nt))) return rc;
  if ((rc = gmi_cilia_compute_velocities(c, nt, its))) return rc;
  if ((rc = gmi_cilia_integrate_bodies(c))) return rc;
  return gmi_cilia_end_of_step(c, nt);
}

int gmi_cilia_get_state(gmi_ctx *c, double *seg_x, double *phase, double *speed, double *body,
                        double *vel_dir) {
  CiliaHost *s = cil_of(c);
  if (!s) return c ? fail(c, GMI_ERR_STATE, "no prescribed-cilia stepper") : GMI_ERR_ARG;
  const CilDev &d = s->d;
  int rc;
  if ((rc = cil_sync(c))) return rc;
  if (seg_x && (rc = soa_to_host(c, d.X, d.lds, d.Ns, seg_x))) return rc;
  if (vel_dir && (rc = soa_to_host(c, d.Kd, d.lds, d.Ns, vel_dir))) return rc;
  if (phase && (rc = d2h(c, phase, d.phase, sizeof(double) * d.Nf)))
    return rc;
  if (speed && (rc = d2h(c, speed, d.speed, sizeof(double) * d.Nf)))
    return rc;
  if (body) {
    std::vector<double> h(size_t(d.Nsw) * CB_REC);
    if ((rc = d2h(c, h.data(), d.body, sizeof(double) * h.size())))
      return rc;
    for (int n = 0; n < d.Nsw; n++)
      for (int k = 0; k < 7; k++) body[7 * n + k] = h[size_t(n) * CB_REC + CB_X + k];
  }
  return GMI_OK;
}

int gmi_cilia_get_solution(gmi_ctx *c, double *f_segs, double *f_blobs, double *v_bodies) {
  CiliaHost *s = cil_of(c);
  if (!s) return c ? fail(c, GMI_ERR_STATE, "no prescribed-cilia stepper") : GMI_ERR_ARG;
  const CilDev &d = s->d;
  int rc;
  if ((rc = cil_sync(c))) return rc;
  if (f_segs && (rc = soa_to_host(c, s->x, d.lds, d.Ns, f_segs))) return rc;
  if (f_blobs && (rc = soa_to_host(c, s->x + d.ob, d.ldb, d.Nb, f_blobs))) return rc;
  if (v_bodies) {
    if (d.free_body) {
      if ((rc = d2h(c, v_bodies, s->x + d.oc, sizeof(double) * 6 * d.Nsw)))
        return rc;
    } else {
      std::vector<double> h(size_t(d.Nsw) * CB_REC);
      if ((rc = d2h(c, h.data(), d.body, sizeof(double) * h.size())))
        return rc;
      for (int n = 0; n < d.Nsw; n++)
        for (int k = 0; k < 6; k++) v_bodies[6 * n + k] = h[size_t(n) * CB_REC + CB_VP + k];
    }
  }
  return GMI_OK;
}

int gmi_cilia_generalised_forces(gmi_ctx *c, double *out) {
  CiliaHost *s = cil_of(c);
  if (!s || !out) return c ? fail(c, GMI_ERR_STATE, "no prescribed-cilia stepper") : GMI_ERR_ARG;
  hipLaunchKernelGGL(k_cf_genforce, dim3(nblk(s->d.Nf, 64)), dim3(64), 0, c->stream, s->d, s->x,
                     s->gout);
  int rc;
  if ((rc = launch_ok(c, "generalised forces"))) return rc;
  return d2h(c, out, s->gout, sizeof(double) * s->d.Nf);
}

int gmi_cilia_segment_velocities(gmi_ctx *c, double *v_segs) {
  CiliaHost *s = cil_of(c);
  if (!s || !v_segs) return c ? fail(c, GMI_ERR_STATE, "no prescribed-cilia stepper") : GMI_ERR_ARG;
  int rc;
  if ((rc = run_pairs(c, kSS | kSB, 0, false, s->x, s->x + s->d.ob, nullptr, nullptr, nullptr,
                      nullptr)))
    return rc;
  return soa_to_host(c, c->Vs, s->d.lds, s->d.Ns, v_segs);
}

int gmi_cilia_system_size(gmi_ctx *c, int *n) {
  CiliaHost *s = cil_of(c);
  if (!s || !n) return c ? fail(c, GMI_ERR_STATE, "no prescribed-cilia stepper") : GMI_ERR_ARG;
  *n = s->nsys_ref;
  return GMI_OK;
}

int gmi_cilia_rhs(gmi_ctx *c, double *out) {
  CiliaHost *s = cil_of(c);
  if (!s || !out) return c ? fail(c, GMI_ERR_STATE, "no prescribed-cilia stepper") : GMI_ERR_ARG;
  hipLaunchKernelGGL(k_cil_rhs, dim3(s->ntiles), dim3(256), 0, c->stream, s->d, s->tiles, s->rhs);
  int rc;
  if ((rc = launch_ok(c, "cilia rhs"))) return rc;
  return dev_to_ref(c, s, s->rhs, out);
}

int gmi_cilia_apply_operator(gmi_ctx *c, const double *in, double *out) {
  CiliaHost *s = cil_of(c);
  if (!s || !in || !out) return c ? fail(c, GMI_ERR_STATE, "no prescribed-cilia stepper") : GMI_ERR_ARG;
  int rc;
  if ((rc = ref_to_dev(c, s, in, s->z))) return rc;
  if ((rc = hip_check(c, hipMemsetAsync(s->tmp, 0, sizeof(double) * s->d.n, c->stream), "zero")))
    return rc;
  CiliaOp op{c, s};
  if ((rc = op.mult(s->z, s->tmp, 0, nullptr, false))) return rc;
  return dev_to_ref(c, s, s->tmp, out);
}

int gmi_cilia_apply_preconditioner(gmi_ctx *c, const double *in, double *out) {
  CiliaHost *s = cil_of(c);
  if (!s || !in || !out) return c ? fail(c, GMI_ERR_STATE, "no prescribed-cilia stepper") : GMI_ERR_ARG;
  int rc;
  if ((rc = ref_to_dev(c, s, in, s->z))) return rc;
  if ((rc = hip_check(c, hipMemsetAsync(s->tmp, 0, sizeof(double) * s->d.n, c->stream), "zero")))
    return rc;
  CiliaOp op{c, s};
  if ((rc = op.precond(s->z, s->tmp, nullptr, -1))) return rc;
  return dev_to_ref(c, s, s->tmp, out);
}

int gmi_cilia_set_precision(gmi_ctx *c, int precision) {
  CiliaHost *s = cil_of(c);
  if (!s) return c ? fail(c, GMI_ERR_STATE, "no prescribed-cilia stepper") : GMI_ERR_ARG;
  if (precision != GMI_FP64 && precision != GMI_FP32_KAHAN)
    return fail(c, GMI_ERR_ARG, "unknown precision");
  if (c->wall && precision != GMI_FP64)
    return fail(c, GMI_ERR_ARG, "the wall-bounded mobility is FP64 only");
  s->precision = precision;
  return GMI_OK;
}

int gmi_cilia_set_gram_schmidt(gmi_ctx *c, int passes) {
  CiliaHost *s = cil_of(c);
  if (!s) return c ? fail(c, GMI_ERR_STATE, "no prescribed-cilia stepper") : GMI_ERR_ARG;
  if (passes != 1 && passes != 2) return fail(c, GMI_ERR_ARG, "passes is 1 or 2");
  int rc;
  if ((rc = cil_sync(c))) return rc;
  s->drop_graphs();
  s->gs_passes = passes;
  return GMI_OK;
}

int gmi_cilia_set_graphs(gmi_ctx *c, int use_graphs, int chunk) {
  CiliaHost *s = cil_of(c);
  if (!s) return c ? fail(c, GMI_ERR_STATE, "no prescribed-cilia stepper") : GMI_ERR_ARG;
  if (chunk < 1 || chunk > 64) return fail(c, GMI_ERR_ARG, "chunk must be in [1, 64]");
  int rc;
  if ((rc = cil_sync(c))) return rc;
  s->drop_graphs();
  s->use_graphs = use_graphs != 0;
  s->chunk = chunk;
  return GMI_OK;
}

}  // extern "C"
