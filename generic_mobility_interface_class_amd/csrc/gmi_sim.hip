// gmi_sim.hip -- device-resident implicit time step (include/gmi_sim.h).
//
// Everything the reference keeps in host objects (segment/filament/body
// state, Broyden C and D, the GMRES Krylov basis) lives in HBM here, in
// structure-of-arrays form; each reference method becomes one or two
// kernels and the host only reads convergence flags.  Decomposition:
//   per segment  : residual rows                      (thread per segment)
//   per filament : initial guess, forces, robot arm,   (thread per filament)
//                  update; Jacobian assembly, Gauss-   (workgroup per filament)
//                  Jordan inverse, Schur blocks, J^-1 x
//   per swimmer  : body state, Schur complement,       (thread / workgroup
//                  preconditioner and K products        per swimmer)
//   per vector   : GMRES Arnoldi, Givens,               (grid-wide kernels)
//                  Broyden rank-one updates
// Numerics follow the reference (file:line on every kernel); GMRES uses
// classical Gram-Schmidt instead of modified Gram-Schmidt (one pass per
// step by default, as the reference; two on request) so that a pass is one
// batched dot-product sweep rather than k dependent ones.
#include "gmi_internal.hpp"
#include "gmi_ctx.hpp"
#include "gmi_lie.hpp"
#include "gmi_sim_state.hpp"
#include "gmi_krylov.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <utility>
#include <vector>

namespace gmi {
namespace {

constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr int kHistMax = 16;  // right-hand sides kept for the fitted GMRES guess

__device__ __forceinline__ d3 rowvec(const double *p) { return {p[0], p[1], p[2]}; }

// -------------------------------------------------------------------------
// small dense helpers
__device__ void rotate66(const double *N, const m3 &Q, double *out) {
  const m3 QT = mtrans(Q);
  for (int bi = 0; bi < 2; bi++)
    for (int bj = 0; bj < 2; bj++) {
      m3 B;
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) B.a[3 * i + j] = N[6 * (3 * bi + i) + 3 * bj + j];
      const m3 C = mm(mm(Q, B), QT);
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) out[6 * (3 * bi + i) + 3 * bj + j] = C.a[3 * i + j];
    }
}

// Gauss-Jordan with partial pivoting of a small row-major matrix (n <= 6).
__device__ bool gj_small(double *A, int n) {
  int piv[6];
  for (int i = 0; i < n; i++) piv[i] = i;
  for (int k = 0; k < n; k++) {
    int p = k;
    double best = fabs(A[k * n + k]);
    for (int i = k + 1; i < n; i++)
      if (fabs(A[i * n + k]) > best) { best = fabs(A[i * n + k]); p = i; }
    if (best == 0.0) return false;
    if (p != k) {
      for (int j = 0; j < n; j++) { const double t = A[k * n + j]; A[k * n + j] = A[p * n + j]; A[p * n + j] = t; }
      const int t = piv[k]; piv[k] = piv[p]; piv[p] = t;
    }
    const double inv = 1.0 / A[k * n + k];
    A[k * n + k] = 1.0;
    for (int j = 0; j < n; j++) A[k * n + j] *= inv;
    for (int i = 0; i < n; i++) {
      if (i == k) continue;
      const double f = A[i * n + k];
      if (f == 0.0) continue;
      A[i * n + k] = 0.0;
      for (int j = 0; j < n; j++) A[i * n + j] -= f * A[k * n + j];
    }
  }
  double B[36];
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) B[i * n + piv[j]] = A[i * n + j];
  for (int k = 0; k < n * n; k++) A[k] = B[k];
  return true;
}

// -------------------------------------------------------------------------
// filament helpers (one thread owns one filament)

__device__ dq seg_q(const SimDev &a, int g) { return ld4(a.Q, a.lds, g); }
__device__ d3 seg_t(const SimDev &a, int g) { return qtangent(seg_q(a, g)); }

// filament.cpp:416-436: positions from the base along the segment
// tangents (tangents of the filament's segments in tn[3 S], e.g. in LDS)
__device__ void robot_arm(const SimDev &a, int g0, const double *tn) {
  d3 x = ld3(a.X, a.lds, g0);
  d3 t1 = ga3(tn);
  for (int j = 1; j < a.S; j++) {
    const d3 t2 = ga3(tn + 3 * j);
    x = x + (0.5 * a.DL) * (t1 + t2);
    st3(a.X, a.lds, g0 + j, x);
    t1 = t2;
  }
}

// segment.cpp:81-90: u += du, q = exp(u) qm1
__device__ void seg_update(const SimDev &a, int g, d3 du) {
  const d3 u = ld3(a.U, a.lds, g) + du;
  st3(a.U, a.lds, g, u);
  st4(a.Q, a.lds, g, qmul(lie_exp(u), ld4(a.Qm1, a.lds, g)));
}

__device__ d3 base_position(const SimDev &a, int f, const double *b) {
  const m3 R = rotm(ga4(b + BQ));
  return ga3(b + BX) + mv(R, rowvec(a.fil_ref + 3 * f));
}

// filament.cpp:438-463 (accept_state_from_rigid_body) without its robot arm:
// the base at the body, segment 0 turned with the body
__device__ void accept_base(const SimDev &a, int f, int g0) {
  const double *b = a.body + size_t(a.fil_sw[f]) * BREC;
  st3(a.X, a.lds, g0, base_position(a, f, b));
  seg_update(a, g0, ga3(b + BU) - ld3(a.U, a.lds, g0));
}

// -------------------------------------------------------------------------
// initial guess (rigid_body.cpp:247-332, segment.cpp:38-79, filament.cpp:825-1018)

__global__ void k_body_initial_guess(SimDev a, int nt) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= a.Nsw) return;
  double *b = a.body + size_t(n) * BREC;
  const d3 x = ga3(b + BX), xm1 = ga3(b + BXM1), xm2 = ga3(b + BXM2);
  const d3 u = ga3(b + BU), um1 = ga3(b + BUM1);
  d3 gx, gu;
  if (nt <= a.neuler) {
    gx = 2.0 * x - xm1;
    gu = u;
  } else {
    gx = 3.0 * (x - xm1) + xm2;
    gu = 2.0 * u - um1;
  }
  sa3(b + BXM2, xm1); sa3(b + BXM1, x); sa3(b + BX, gx);
  sa3(b + BUM1, u); sa3(b + BU, gu);
  const dq qm1 = ga4(b + BQ);
  sa4(b + BQM1, qm1);
  const dq q = qmul(lie_exp(gu), qm1);
  sa4(b + BQ, q);
  const m3 R = rotm(q);
  for (int k = 0; k < 9; k++) b[BQI + k] = R.a[k];
}

// element k of a [value | m1 | m2] record of n-vectors: the linear (second
// step) or quadratic extrapolation of the value, history shifted
__device__ void extrap_one(double *v, int n, int k, int nt) {
  double &x = v[k], &m1 = v[n + k], &m2 = v[2 * n + k];
  if (nt == 1) {
    m1 = x;
  } else if (nt == 2) {
    m2 = m1; m1 = x; x = 2.0 * m1 - m2;
  } else if (nt > 2) {
    const double t = 3.0 * (x - m1) + m2;
    m2 = m1; m1 = x; x = t;
  }
}

constexpr int kMaxSeg = 64;    // segments per filament of the time step
constexpr int kFilLanes = 64;  // one wave per filament (NSEG <= kMaxSeg)

// One wave per filament: lane j shifts and extrapolates segment j's state
// and updates its frame; lane 0 then takes the base from the body
// (accept_base) and lays the robot arm out in segment order from the
// tangents in LDS; the multiplier extrapolations are elementwise across the
// lanes.
__global__ __launch_bounds__(kFilLanes) void k_fil_initial_guess(SimDev a, int nt) {
  __shared__ double stn[3 * kFilLanes];
  const int f = blockIdx.x, j = threadIdx.x;
  const int S = a.S, g0 = f * S;
  if (j < S) {
    const int g = g0 + j;
    const d3 x = ld3(a.X, a.lds, g);
    st3(a.Xm2, a.lds, g, ld3(a.Xm1, a.lds, g));
    st3(a.Xm1, a.lds, g, x);
    d3 u = ld3(a.U, a.lds, g);
    if (nt <= a.neuler) {
      st3(a.Um1, a.lds, g, u);
    } else {
      const d3 t = 2.0 * u - ld3(a.Um1, a.lds, g);
      st3(a.Um1, a.lds, g, u);
      u = t;
      st3(a.U, a.lds, g, u);
    }
    const dq q = seg_q(a, g);
    st4(a.Qm1, a.lds, g, q);
    st4(a.Q, a.lds, g, qmul(lie_exp(u), q));
  }
  if (j == 0) accept_base(a, f, g0);
  if (j < S) sa3(stn + 3 * j, seg_t(a, g0 + j));
  __syncthreads();
  if (j == 0) robot_arm(a, g0, stn);
  // lambda: [value | m1 | m2] blocks of 3(S-1) inside a 9S record
  const int nl = 3 * (S - 1);
  double *L = a.lam + size_t(f) * 9 * S;
  for (int k = j; k < nl; k += kFilLanes) extrap_one(L, nl, k, nt);
  if (j < 3) extrap_one(a.teth + 9 * f, 3, j, nt);
  else if (j < 6) extrap_one(a.clamp + 9 * f, 3, j - 3, nt);
}

// -------------------------------------------------------------------------
// forces (filament.cpp:1096-1220, swimmer.cpp:276-325)

// body-frame moment of link i (filament.cpp:1096-1110)
__device__ d3 bfm(const SimDev &a, dq q1, dq q2, d3 st) {
  const dq qm = qconj(qmid(q1, q2));
  const double s = 1.0 / a.DL;
  const dq dqs = {(q2.s - q1.s) * s, (q2.x - q1.x) * s, (q2.y - q1.y) * s, (q2.z - q1.z) * s};
  const dq k = qmul(qm, dqs);
  return {a.kt * (2.0 * k.x - st.x), a.kb * (2.0 * k.y - st.y), a.kb * (2.0 * k.z - st.z)};
}

__device__ d3 strain_twist(const SimDev &a, int f, double t) {
  return {0.0, 0.0, a.camp * sin(2.0 * kPi * t / a.cper + a.phase[f])};
}

// One wave per filament (NSEG <= 64): lane j owns segment j and the
// link (j, j+1).  The link moments run in parallel; each segment's force and
// torque is assembled from its two links in the order of the sequential loop
// over links (the same operations, so the same rounding), and lane 0 sums
// the filament's share of the body balance in segment order.
// The instability-driven build adds the follower force -efm t at the tip
// and, when `perturb` (nt = 0), the two perturbation forces on segments 4 and
// 12, after the link terms (filament.cpp:1177-1199).
__global__ __launch_bounds__(kFilLanes) void k_fil_forces(SimDev a, int perturb) {
  const double t = a.step_t[0];
  __shared__ double sl[3 * kFilLanes], sM[3 * kFilLanes], sF[6 * kFilLanes];
  const int f = blockIdx.x, j = threadIdx.x;
  const int S = a.S, g0 = f * S, ld = a.lds;
  const double *L = a.lam + size_t(f) * 9 * S;
  double *F = a.Fint;
  dq qj = {1.0, 0.0, 0.0, 0.0};
  d3 tj = {0, 0, 0}, lam = {0, 0, 0}, M = {0, 0, 0};
  if (j < S) {
    qj = seg_q(a, g0 + j);
    tj = qtangent(qj);
  }
  if (j < S - 1) {
    const d3 st = strain_twist(a, f, t);
    lam = rowvec(L + 3 * j);
    const dq qn = seg_q(a, g0 + j + 1);
    M = mv(rotm(qmid(qj, qn)), bfm(a, qj, qn, st));
    sa3(sl + 3 * j, lam);
    sa3(sM + 3 * j, M);
  }
  __syncthreads();
  if (j < S) {
    // link j-1 sets segment j, link j then adds its terms
    d3 Fj = {0, 0, 0}, Tj = {0, 0, 0};
    if (j > 0) {
      const d3 lp = ga3(sl + 3 * (j - 1));
      Fj = lp;
      Tj = (-0.5 * a.DL) * cross3(tj, lp);
      Tj = Tj - ga3(sM + 3 * (j - 1));
    }
    if (j < S - 1) {
      Fj = Fj - lam;
      Tj = Tj - (0.5 * a.DL) * cross3(tj, lam);
      Tj = Tj + M;
    }
    if (j == 0) Fj = Fj + rowvec(a.teth + 9 * f);
    if (a.efm != 0.0) {
      if (j == S - 1) Fj = Fj - a.efm * tj;
      if (perturb) {
        if (j == 4) Fj = Fj + rowvec(a.pert + 6 * f);
        if (j == 12) Fj = Fj + rowvec(a.pert + 6 * f + 3);
      }
    }
    if (j == 0) Tj = Tj + dexpinv_T(ld3(a.U, ld, g0), rowvec(a.clamp + 9 * f));
    const int g = g0 + j;
    F[g] = Fj.x; F[ld + g] = Fj.y; F[2 * ld + g] = Fj.z;
    F[3 * ld + g] = Tj.x; F[4 * ld + g] = Tj.y; F[5 * ld + g] = Tj.z;
    sa3(sF + 6 * j, Fj);
    sa3(sF + 6 * j + 3, Tj);
  }
  __syncthreads();
  if (j != 0) return;
  // this filament's share of the free-swimming balance
  const double *b = a.body + size_t(a.fil_sw[f]) * BREC;
  const d3 bx = ga3(b + BX);
  d3 sf = {0, 0, 0}, stq = {0, 0, 0};
  for (int k = 0; k < S; k++) {
    const d3 fs = ga3(sF + 6 * k), ts = ga3(sF + 6 * k + 3);
    sf = sf + fs;
    stq = stq + ts + cross3(ld3(a.X, ld, g0 + k) - bx, fs);
  }
  double *o = a.fil_bf + 6 * f;
  o[0] = sf.x; o[1] = sf.y; o[2] = sf.z; o[3] = stq.x; o[4] = stq.y; o[5] = stq.z;
}

// Sums v[0..NV) over the 256 threads of a workgroup into out[] of thread 0
// (waves first, then the four wave sums in wave order).  sh: 4 NV doubles.
template <int NV>
__device__ void wg_sums(const double *v, double *out, double *sh) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int k = 0; k < NV; k++) {
    const double t = wave_sum(v[k]);
    if (lane == 0) sh[4 * k + w] = t;
  }
  __syncthreads();
  if (threadIdx.x == 0)
    for (int k = 0; k < NV; k++) out[k] = ((sh[4 * k] + sh[4 * k + 1]) + sh[4 * k + 2]) + sh[4 * k + 3];
  __syncthreads();
}

// One workgroup of 256 per swimmer.
__global__ __launch_bounds__(256) void k_body_forces(SimDev a) {
  __shared__ double sh[4 * 6];
  const int n = blockIdx.x;
  double *b = a.body + size_t(n) * BREC;
  double v[6] = {0, 0, 0, 0, 0, 0};
  const int f0 = a.sw_fil0[n], f1 = f0 + a.sw_nfil[n];
  for (int f = f0 + threadIdx.x; f < f1; f += 256)
    for (int k = 0; k < 6; k++) v[k] -= a.fil_bf[6 * f + k];
  double s[6];
  wg_sums<6>(v, s, sh);
  if (threadIdx.x == 0)
    for (int k = 0; k < 6; k++) b[BF + k] = s[k];
}

// -------------------------------------------------------------------------
// compute_velocities pieces (mobility_solver.cpp:132-192, 361-421)

__global__ void k_blob_positions(SimDev a) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= a.Nb) return;
  const double *b = a.body + size_t(a.blob_sw[m]) * BREC;
  const m3 R = rotm(ga4(b + BQ));
  st3(a.Xb, a.ldb, m, ga3(b + BX) + mv(R, ld3(a.blob_ref, a.ldb, m)));
}

// swimmer.f += sum over its blobs of (F_rep, r x F_rep)
__global__ __launch_bounds__(256) void k_body_repulsion(SimDev a) {
  __shared__ double sh[256];
  const int n = blockIdx.x;
  double *b = a.body + size_t(n) * BREC;
  const m3 R = rotm(ga4(b + BQ));
  const int m0 = a.sw_blob0[n], m1 = m0 + a.sw_nblob[n];
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int m = m0 + threadIdx.x; m < m1; m += 256) {
    const d3 fr = ld3(a.Frep, a.ldb, m);
    const d3 c = cross3(mv(R, ld3(a.blob_ref, a.ldb, m)), fr);
    acc[0] += fr.x; acc[1] += fr.y; acc[2] += fr.z;
    acc[3] += c.x; acc[4] += c.y; acc[5] += c.z;
  }
  for (int k = 0; k < 6; k++) {
    const double s = block_sum<256>(acc[k], sh);
    if (threadIdx.x == 0) b[BF + k] += s;
  }
}

// rhs = [-v_blobs ; -swimmer.f]  (GMRES vector: lambda SoA [3][ldb], U [6 Nsw])
__global__ void k_assemble_rhs(SimDev a, double *rhs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 3 * a.ldb) {
    const int m = i % a.ldb;
    rhs[i] = (m < a.Nb) ? -a.Vb[i] : 0.0;
  } else if (i < a.npad) {
    const int r = i - 3 * a.ldb;
    rhs[i] = -a.body[size_t(r / 6) * BREC + BF + r % 6];
  }
}

// Blob tiles: up to 256 consecutive blobs of one swimmer (tile table built
// at creation; the same tiles as the blob-blob pair work list), so the
// per-blob parts of the preconditioner and of the K products run on many
// workgroups.  Per-swimmer sums over tiles are taken in tile order by each
// consumer (deterministic, no cross-workgroup synchronisation inside a
// kernel).

// -(sum over the tile's blobs of (v_m, d_m x v_m)) = the tile's share of -K^T v
// (wave sums, then the 4 waves in order; sh holds 24 doubles)
__device__ void tile_kt_vals(const m3 &Qm, const SimDev &a, int m, bool live, d3 l,
                             double *part, double *sh) {
  double acc[6] = {0, 0, 0, 0, 0, 0};
  if (live) {
    const d3 c = cross3(mv(Qm, ld3(a.blob_ref, a.ldb, m)), l);
    acc[0] = -l.x; acc[1] = -l.y; acc[2] = -l.z;
    acc[3] = -c.x; acc[4] = -c.y; acc[5] = -c.z;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int k = 0; k < 6; k++) {
    const double r = wave_sum(acc[k]);
    if (lane == 0) sh[6 * wave + k] = r;
  }
  __syncthreads();
  if (threadIdx.x < 6)
    part[threadIdx.x] = ((sh[threadIdx.x] + sh[6 + threadIdx.x]) + sh[12 + threadIdx.x]) +
                        sh[18 + threadIdx.x];
}

// Rotated body matrices of every swimmer for one rigid-body solve (the body
// orientation is fixed during it): rot[n] = [Q N_ref Q^T | Q (K^T K)^-1 Q^T]
__global__ void k_rotate_body(SimDev a, double *rot) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= a.Nsw) return;
  const double *b = a.body + size_t(n) * BREC;
  const m3 Qm = rotm(ga4(b + BQ));
  rotate66(b + BN, Qm, rot + 72 * size_t(n));
  rotate66(b + BK, Qm, rot + 72 * size_t(n) + 36);
}

// Preconditioner pass 1 (mobility_solver.cpp:651-717): the tile's share of
// -K^T in.  `in` is the Krylov vector of the previous Arnoldi step, still
// unnormalised: its norm scal[3] (from k_givens) is divided out here and the
// normalised vector written back (each tile owns its blob rows, the
// swimmer's first tile its 6 body rows).  norm_slot < 0: `in` is used as is.
__global__ __launch_bounds__(256) void k_precond_kt(SimDev a, const BTile *tiles,
                                                    const int *sw_tile0, double *in,
                                                    const double *scal, int norm_slot,
                                                    double *part, const int *flags) {
  if (flags && flags[F_DONE]) return;
  __shared__ double sh[256];
  const BTile T = tiles[blockIdx.x];
  const double *b = a.body + size_t(T.sw) * BREC;
  const m3 Qm = rotm(ga4(b + BQ));
  const double nrm = norm_slot >= 0 ? scal[norm_slot] : 1.0;
  const bool live = threadIdx.x < T.n;
  const int m = T.m0 + threadIdx.x;
  d3 l = {0, 0, 0};
  if (live) {
    l = ld3(in, a.ldb, m);
    if (norm_slot >= 0) {
      l = {l.x / nrm, l.y / nrm, l.z / nrm};
      st3(in, a.ldb, m, l);
    }
  }
  if (norm_slot >= 0 && blockIdx.x == sw_tile0[T.sw] && threadIdx.x < 6)
    in[3 * a.ldb + 6 * T.sw + threadIdx.x] /= nrm;
  tile_kt_vals(Qm, a, m, live, l, part + 6 * blockIdx.x, sh);
}

// Preconditioner pass 2: each tile sums its swimmer's -K^T in partials (tile
// order), solves the body rows out_U = (K^T K)^-1 b - N in_U, and writes its
// blobs out = M^-1 (in + U + W x d); the swimmer's first tile writes out_U.
// Also the tile's share of -K^T out, which the K products need next.
__global__ __launch_bounds__(256) void k_precond_blobs(SimDev a, const BTile *tiles,
                                                       const int *sw_tile0, const int *sw_ntile,
                                                       const double *part, const double *rot,
                                                       const double *in, double *out,
                                                       double *part_out, const int *flags) {
  if (flags && flags[F_DONE]) return;
  __shared__ double sh[256];
  __shared__ double sb[6], so[6];
  const BTile T = tiles[blockIdx.x];
  const int n = T.sw, t = threadIdx.x;
  const int vid = 3 * a.ldb + 6 * n;
  if (t < 6) {
    double s = 0.0;
    for (int q = sw_tile0[n]; q < sw_tile0[n] + sw_ntile[n]; q++) s += part[6 * q + t];
    sb[t] = s;
  }
  __syncthreads();
  if (t < 6) {
    const double *bm = rot + 72 * size_t(n), *kk = bm + 36;
    double s1 = 0, s2 = 0;
    for (int k = 0; k < 6; k++) { s1 += kk[6 * t + k] * sb[k]; s2 += bm[6 * t + k] * in[vid + k]; }
    so[t] = s1 - s2;
    if (blockIdx.x == sw_tile0[n]) out[vid + t] = s1 - s2;
  }
  __syncthreads();
  const double *b = a.body + size_t(n) * BREC;
  const m3 Qm = rotm(ga4(b + BQ));
  const bool live = t < T.n;
  const int m = T.m0 + t;
  d3 z = {0, 0, 0};
  if (live) {
    const d3 V = {so[0], so[1], so[2]}, W = {so[3], so[4], so[5]};
    const double Minv = 6.0 * kPi * a.mu * a.ab;
    const d3 d = mv(Qm, ld3(a.blob_ref, a.ldb, m));
    z = Minv * (ld3(in, a.ldb, m) + V + cross3(W, d));
    st3(out, a.ldb, m, z);
  }
  tile_kt_vals(Qm, a, m, live, z, part_out + 6 * blockIdx.x, sh);
}

// system_matrix_mult K products (mobility_solver.cpp:905-964): blob rows
// out = M_bb z - (U + W x d), with M_bb z summed here from the pair kernel's
// split partials (split order) when the tile was split; body rows
// out_U = -K^T z from the tile shares of k_precond_blobs (tile order),
// written by the swimmer's first tile.
__global__ __launch_bounds__(256) void k_kprod(SimDev a, const BTile *tiles, const int *sw_tile0,
                                               const int *sw_ntile, const TileDesc *ptiles,
                                               const double *ppart, const double *z,
                                               const double *vb, double *out,
                                               const double *zpart, const int *flags) {
  if (flags && flags[F_DONE]) return;
  const BTile T = tiles[blockIdx.x];
  const int n = T.sw, t = threadIdx.x;
  const int vid = 3 * a.ldb + 6 * n;
  if (blockIdx.x == sw_tile0[n] && t < 6) {
    double s = 0.0;
    for (int q = sw_tile0[n]; q < sw_tile0[n] + sw_ntile[n]; q++) s += zpart[6 * q + t];
    out[vid + t] = s;
  }
  if (t >= T.n) return;
  const int m = T.m0 + t;
  const TileDesc P = ptiles[blockIdx.x];
  d3 v;
  if (P.nsplit == 1) {
    v = ld3(vb, a.ldb, m);
  } else {
    double s[3] = {0.0, 0.0, 0.0};
    const double *pp = ppart + size_t(P.part0) * 6 * kBlock + t;
#pragma unroll 4
    for (int q = 0; q < P.nsplit; ++q)
      for (int c = 0; c < 3; ++c) s[c] += pp[size_t(q) * 6 * kBlock + c * kBlock];
    v = {s[0], s[1], s[2]};
  }
  const double *b = a.body + size_t(n) * BREC;
  const m3 Qm = rotm(ga4(b + BQ));
  const d3 V = {z[vid], z[vid + 1], z[vid + 2]}, W = {z[vid + 3], z[vid + 4], z[vid + 5]};
  const d3 d = mv(Qm, ld3(a.blob_ref, a.ldb, m));
  st3(out, a.ldb, m, v - (V + cross3(W, d)));
}

// ---- fused Arnoldi step of the rigid-body solve ----------------------------
// With the reference's preconditioner an Arnoldi step runs on a chunk
// partition of the GMRES vector: workgroup b owns chunk b's blob rows (three
// components of up to 64 blobs of one swimmer) and, if it is its swimmer's
// first chunk, the swimmer's 6 body rows.  Five launches per step instead of
// eight:
//   k_arnoldi_precond  Givens step of the previous iteration (every
//                      workgroup takes it, workgroup 0 stores it), the
//                      previous vector's normalisation, the preconditioner
//   k_pair (bb)        the operator's pair sums (unchanged)
//   k_kprod_dots       K products, then the first Gram-Schmidt pass's dots
//   k_gs_chunk<1>      h1, w -= Q h1, the second pass's dots
//   k_gs_chunk<2>      h2, w -= Q h2, |w|^2 shares and the -K^T w shares the
//                      next preconditioner needs
// With one Gram-Schmidt pass per step (the default; gmi_sim_set_gram_schmidt)
// k_gs_chunk<2> takes the first pass's dots and k_gs_chunk<1> is not launched.
// Partials are per chunk and every consumer sums them in chunk order.  Rows
// past a chunk's blobs (the padding of the SoA layout) are zero in every
// Krylov vector and take no part.  Thread t < 198 owns slot t of its chunk:
// slots 0..191 are component t / 64 of blob t % 64, slots 192..197 body rows.

constexpr int kChunk = 64;                 // blobs per chunk
constexpr int kWideChunks = 256;           // above: column partials summed once per column
constexpr int kChunkRows = 3 * kChunk + 6; // slots per chunk

// GMRES row of slot q of chunk T, or -1 when the chunk has no such row.
__device__ __forceinline__ int chunk_row(const SimDev &a, const BTile &T, bool first, int q) {
  if (q < 3 * kChunk) {
    const int i = q % kChunk;
    return i < T.n ? (q / kChunk) * a.ldb + T.m0 + i : -1;
  }
  return first && q < kChunkRows ? 3 * a.ldb + 6 * T.sw + (q - 3 * kChunk) : -1;
}

// The fused Gram-Schmidt kernels run 1024 threads (16 waves) per chunk: a
// step's columns are spread over the waves so that each phase is about one
// round trip to memory.
constexpr int kGsThreads = 1024;
constexpr int kGsWaves = kGsThreads / 64;

// Dots of the chunk's rows (values sv by slot) with columns j < k of Qb:
// wave w takes the columns j = w (mod kGsWaves), kDotBatch at a time, and
// writes part[j * nparts + chunk].
__device__ void chunk_col_dots(const SimDev &a, const BTile &T, bool first, const double *sv,
                               const double *Qb, size_t ldq, int k, double *part, int nparts) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  constexpr int NR = (kChunkRows + 63) / 64;
  int rows[NR];
  double vals[NR];
#pragma unroll
  for (int r = 0; r < NR; r++) {
    const int q = lane + 64 * r;
    rows[r] = chunk_row(a, T, first, q);
    vals[r] = rows[r] >= 0 ? sv[q] : 0.0;
  }
  for (int j0 = wave; j0 < k; j0 += kGsWaves * kDotBatch) {
    double acc[kDotBatch];
#pragma unroll
    for (int c = 0; c < kDotBatch; c++) acc[c] = 0.0;
#pragma unroll
    for (int r = 0; r < NR; r++) {
      if (rows[r] < 0) continue;
#pragma unroll
      for (int c = 0; c < kDotBatch; c++) {
        // past the last column a batch re-reads column k-1 and stores nothing
        const int j = min(j0 + kGsWaves * c, k - 1);
        acc[c] += Qb[ldq * j + rows[r]] * vals[r];
      }
    }
#pragma unroll
    for (int c = 0; c < kDotBatch; c++) {
      const double s = wave_sum(acc[c]);
      const int j = j0 + kGsWaves * c;
      if (lane == 0 && j < k) part[size_t(j) * nparts + blockIdx.x] = s;
    }
  }
}

// h[j] = sum over chunks of part[j * nparts + q] into LDS, and from
// workgroup 0 into hout.  Wave w takes the columns j = w (mod kGsWaves),
// four at a time; lane l sums the chunks q = l (mod 64), then the wave sums
// the lanes (a fixed order, the same in every workgroup).
__device__ void chunk_col_sums(const double *part, int k, int nparts, double *hs, double *hout) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int j0 = wave; j0 < k; j0 += 4 * kGsWaves) {
    double acc[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const int j = min(j0 + kGsWaves * c, k - 1);
      const double *p = part + size_t(j) * nparts;
      double s = 0.0;
      for (int q = lane; q < nparts; q += 64) s += p[q];
      acc[c] = s;
    }
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const double s = wave_sum(acc[c]);
      const int j = j0 + kGsWaves * c;
      if (lane == 0 && j < k) {
        hs[j] = s;
        if (blockIdx.x == 0) hout[j] = s;
      }
    }
  }
}

// Sum over the swimmer's chunks [q0, q0 + nq) of part[6 q + c] for c < 6,
// into out[c]: lanes 32 c' + l of waves 0..2 (c = 2 wave + c') take the
// chunks q0 + l (mod 32), then each half-wave sums its lanes.  Every thread
// of the workgroup calls it; out is LDS.
__device__ void swimmer_share_sums(const double *part, int q0, int nq, double *out) {
  const int t = threadIdx.x;
  double s = 0.0;
  const int c = t >> 5, l = t & 31;
  if (c < 6)
    for (int q = l; q < nq; q += 32) s += part[6 * (q0 + q) + c];
  s = half_wave_sum(s);
  if (c < 6 && l == 0) out[c] = s;
  __syncthreads();
}

// The chunk's share of -K^T w from its blob rows (values l of blob t in
// wave 0): one wave sum per component.
__device__ void chunk_kt_vals(const m3 &Qm, const SimDev &a, int m, bool live, d3 l,
                              double *part) {
  if (threadIdx.x >= 64) return;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  if (live) {
    const d3 c = cross3(mv(Qm, ld3(a.blob_ref, a.ldb, m)), l);
    acc[0] = -l.x; acc[1] = -l.y; acc[2] = -l.z;
    acc[3] = -c.x; acc[4] = -c.y; acc[5] = -c.z;
  }
#pragma unroll
  for (int k = 0; k < 6; k++) {
    const double r = wave_sum(acc[k]);
    if (threadIdx.x == 0) part[k] = r;
  }
}

// Preconditioner of Arnoldi step it (mobility_solver.cpp:651-717) on the
// previous vector w = in.  qn_mode 1: first the Givens step of iteration
// it-1 from the previous step's shares (|w|^2 in npart, -K^T w in kpart),
// taken by every workgroup and stored by workgroup 0, then w is divided by
// its norm; 0: a k_givens launch took that step and left |w| in scal[3];
// 2: w is normalised already (the first step).  The workgroup writes its
// blob rows of w / |w| back (the body rows are normalised by the swimmer's
// first chunk in k_kprod_dots: every chunk of the swimmer reads them here),
// and out = P^-1 (w / |w|) with the chunk's -K^T out share.
__global__ __launch_bounds__(256) void k_arnoldi_precond(
    SimDev a, const BTile *chunks, const int *sw_chunk0, const int *sw_nchunk,
    const double *kpart, const double *npart, int nparts, const double *rot, double *in,
    double *out, double *part_out, GivensArgs g, int it_prev, int qn_mode) {
  if (g.flags[F_DONE]) return;
  __shared__ double sh[256];
  __shared__ double hc[kMaxCols + 1], scs[kMaxCols], ssn[kMaxCols];
  __shared__ int sconv;
  __shared__ double sb[6], so[6];
  double qn = 1.0;
  if (qn_mode == 1) {
    qn = norm_of_parts(npart, nparts, sh);
    if (givens_column(qn, it_prev, g, blockIdx.x == 0, hc, scs, ssn, &sconv)) return;
  } else if (qn_mode == 0) {
    qn = g.scal[3];
  }
  const BTile T = chunks[blockIdx.x];
  const int n = T.sw, t = threadIdx.x;
  const int vid = 3 * a.ldb + 6 * n;
  swimmer_share_sums(kpart, sw_chunk0[n], sw_nchunk[n], sb);
  if (t < 6) sb[t] /= qn;
  __syncthreads();
  if (t < 6) {
    const double *bm = rot + 72 * size_t(n), *kk = bm + 36;
    double s1 = 0, s2 = 0;
    for (int k = 0; k < 6; k++) { s1 += kk[6 * t + k] * sb[k]; s2 += bm[6 * t + k] * (in[vid + k] / qn); }
    so[t] = s1 - s2;
    if (blockIdx.x == sw_chunk0[n]) out[vid + t] = s1 - s2;
  }
  __syncthreads();
  const double *b = a.body + size_t(n) * BREC;
  const m3 Qm = rotm(ga4(b + BQ));
  const bool live = t < T.n;
  const int m = T.m0 + t;
  d3 z = {0, 0, 0};
  if (live) {
    d3 l = ld3(in, a.ldb, m);
    if (qn_mode != 2) {
      l = {l.x / qn, l.y / qn, l.z / qn};
      st3(in, a.ldb, m, l);
    }
    const d3 V = {so[0], so[1], so[2]}, W = {so[3], so[4], so[5]};
    const double Minv = 6.0 * kPi * a.mu * a.ab;
    const d3 d = mv(Qm, ld3(a.blob_ref, a.ldb, m));
    z = Minv * (l + V + cross3(W, d));
    st3(out, a.ldb, m, z);
  }
  tile_kt_vals(Qm, a, m, live, z, part_out + 6 * blockIdx.x, sh);
}

// k_kprod followed by the first Gram-Schmidt pass's dots of the new vector
// out with the columns j < k of Qb.  The blob-blob pair partials of chunk b
// are those of blob tile ch_tile[b].  norm_prev: column k-1's body rows are
// still unnormalised (k_arnoldi_precond left them); the swimmer's first
// chunk divides them by scal[3] before they are used.
__global__ __launch_bounds__(kGsThreads) void k_kprod_dots(
    SimDev a, const BTile *chunks, const int *ch_tile, const int *sw_chunk0,
    const int *sw_nchunk, const TileDesc *ptiles, const double *ppart, const double *z,
    const double *vb, double *out, const double *zpart, const int *flags, double *Qb,
    size_t ldq, int k, int norm_prev, const double *scal, double *part, int nparts) {
  if (flags[F_DONE]) return;
  __shared__ double sv[kChunkRows];
  __shared__ double sz[6], spp[4][3][kChunk];
  const BTile T = chunks[blockIdx.x];
  const int n = T.sw, t = threadIdx.x;
  const int vid = 3 * a.ldb + 6 * n;
  const bool first = blockIdx.x == sw_chunk0[n];
  if (t < kChunkRows) sv[t] = 0.0;
  if (first) swimmer_share_sums(zpart, sw_chunk0[n], sw_nchunk[n], sz);
  if (first && t < 6) {
    if (norm_prev) Qb[ldq * (k - 1) + vid + t] /= scal[3];
    out[vid + t] = sz[t];
    sv[3 * kChunk + t] = sz[t];
  }
  // the pair kernel's split partials: thread (blob i, quarter r) sums the
  // splits q = r (mod 4), then the quarters are added in order
  const TileDesc P = ptiles[ch_tile[blockIdx.x]];
  {
    const int i = t & (kChunk - 1), r = t >> 6;
    if (r < 4 && P.nsplit > 1 && i < T.n) {
      const double *pp = ppart + size_t(P.part0) * 6 * kBlock + (T.m0 + i - P.t0);
      double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll 8
      for (int q = r; q < P.nsplit; q += 4) {
        const double *pq = pp + size_t(q) * 6 * kBlock;
        s0 += pq[0]; s1 += pq[kBlock]; s2 += pq[2 * kBlock];
      }
      spp[r][0][i] = s0; spp[r][1][i] = s1; spp[r][2][i] = s2;
    }
  }
  __syncthreads();
  if (t < T.n) {
    const int m = T.m0 + t;
    d3 v;
    if (P.nsplit == 1) {
      v = ld3(vb, a.ldb, m);
    } else {
      double s[3];
      for (int c = 0; c < 3; c++) s[c] = (spp[0][c][t] + spp[1][c][t]) + (spp[2][c][t] + spp[3][c][t]);
      v = {s[0], s[1], s[2]};
    }
    const double *b = a.body + size_t(n) * BREC;
    const m3 Qm = rotm(ga4(b + BQ));
    const d3 V = {z[vid], z[vid + 1], z[vid + 2]}, W = {z[vid + 3], z[vid + 4], z[vid + 5]};
    const d3 d = mv(Qm, ld3(a.blob_ref, a.ldb, m));
    const d3 w = v - (V + cross3(W, d));
    st3(out, a.ldb, m, w);
    sv[t] = w.x; sv[kChunk + t] = w.y; sv[2 * kChunk + t] = w.z;
  }
  __syncthreads();
  chunk_col_dots(a, T, first, sv, Qb, ldq, k, part, nparts);
}

// Gram-Schmidt pass PASS on the chunk partition: h = sum of the pass's dot
// partials (workgroup 0 stores it in hout), w -= Q h on the chunk's rows
// (thread t its slot t); PASS 1: the dots of the new w with the columns
// (the next pass's partials); PASS 2: the chunk's share of |w|^2 (npart)
// and of -K^T w (kpart).  With one pass per step PASS 2 follows the first
// dots directly and workgroup 0 zeroes hzero, the unused second pass's
// coefficients that the Givens step adds.
template <int PASS>
__global__ __launch_bounds__(kGsThreads) void k_gs_chunk(SimDev a, const BTile *chunks,
                                                         const int *sw_chunk0, const double *Qb,
                                                         size_t ldq, int k, const double *part_in,
                                                         double *hout, double *w,
                                                         double *part_out, int nparts,
                                                         double *npart, double *kpart,
                                                         const int *flags, int h_ready,
                                                         double *hzero) {
  if (flags[F_DONE]) return;
  if (hzero && blockIdx.x == 0)
    for (int j = threadIdx.x; j < k; j += kGsThreads) hzero[j] = 0.0;
  __shared__ double hs[kMaxCols + 16];
  __shared__ double sv[kChunkRows];
  __shared__ double sq[4][256];
  __shared__ double sh[kGsThreads];
  const BTile T = chunks[blockIdx.x];
  const int t = threadIdx.x;
  const bool first = blockIdx.x == sw_chunk0[T.sw];
  if (h_ready)  // many chunks: a k_colsum_wide launch summed the partials once
    for (int j = t; j < k; j += kGsThreads) hs[j] = hout[j];
  else
    chunk_col_sums(part_in, k, nparts, hs, hout);
  if (t < 16) hs[k + t] = 0.0;  // the padding of the last group of 16
  __syncthreads();
  // thread (slot t % 256, quarter t / 256): the quarter's share of
  // sum_j h_j Q_j[row] over the columns j = 16 g + u, g = quarter (mod 4)
  const int slot = t & 255, quarter = t >> 8;
  const int row = chunk_row(a, T, first, slot);
  {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    if (row >= 0)
      for (int j = 16 * quarter; j < k; j += 64) {
        double q[16];
#pragma unroll
        for (int u = 0; u < 16; u++) q[u] = Qb[ldq * min(j + u, k - 1) + row];
#pragma unroll
        for (int u = 0; u < 16; u += 4) {
          s0 += hs[j + u] * q[u];
          s1 += hs[j + u + 1] * q[u + 1];
          s2 += hs[j + u + 2] * q[u + 2];
          s3 += hs[j + u + 3] * q[u + 3];
        }
      }
    sq[quarter][slot] = (s0 + s1) + (s2 + s3);
  }
  __syncthreads();
  double v = 0.0;
  if (t < 256) {
    if (row >= 0) {
      v = w[row] - ((sq[0][t] + sq[1][t]) + (sq[2][t] + sq[3][t]));
      w[row] = v;
    }
    if (t < kChunkRows) sv[t] = v;
  }
  __syncthreads();
  if (PASS == 1) {
    chunk_col_dots(a, T, first, sv, Qb, ldq, k, part_out, nparts);
  } else {
    const double s = block_sum<kGsThreads>(v * v, sh);
    if (t == 0) npart[blockIdx.x] = s;
    const d3 wl = {sv[t % kChunk], sv[kChunk + t % kChunk], sv[2 * kChunk + t % kChunk]};
    const m3 Qm = rotm(ga4(a.body + size_t(T.sw) * BREC + BQ));
    chunk_kt_vals(Qm, a, T.m0 + t, t < T.n, wl, kpart + 6 * blockIdx.x);
  }
}

// -------------------------------------------------------------------------
// vector kernels (GMRES / Broyden)

// ---- history-fitted GMRES initial guess ------------------------------------
// In the body frame the rigid-body system matrix is fixed: blob positions
// and K depend only on the body-frame references.  A new right-hand side,
// rotated into the body frame, is fitted by least squares onto the last m
// right-hand sides of the swimmer, and the same combination of their
// solutions, rotated back, is the GMRES initial guess.  Vectors keep the
// system's layout (blob rows SoA [3][ldb], then 6 body rows per swimmer);
// slot i of the history is Hb + i*ldh (right-hand side) and Hx + i*ldh
// (solution).

// Per tile: the body-frame right-hand side into bnew, and the tile's shares
// of bnew . Hb[i] (i < count) and of bnew . bnew (slot kHistMax).
__global__ __launch_bounds__(256) void k_hist_dots(SimDev a, const BTile *tiles,
                                                   const int *sw_tile0, const double *rhs,
                                                   double *bnew, const double *Hb, size_t ldh,
                                                   const int *hst, double *dpart) {
  __shared__ double sh[4 * (kHistMax + 1)];
  const int count = hst[0];
  const BTile T = tiles[blockIdx.x];
  const int t = threadIdx.x;
  const m3 Qm = rotm(ga4(a.body + size_t(T.sw) * BREC + BQ));
  double acc[kHistMax + 1];
#pragma unroll
  for (int i = 0; i <= kHistMax; i++) acc[i] = 0.0;
  auto add = [&](const d3 v, int o, int ld) {  // o: row offset, ld: component stride
#pragma unroll
    for (int i = 0; i < kHistMax; i++)
      if (i < count) {
        const double *h = Hb + ldh * i + o;
        acc[i] += v.x * h[0] + v.y * h[ld] + v.z * h[2 * ld];
      }
    acc[kHistMax] += dot3(v, v);
  };
  if (t < T.n) {
    const int m = T.m0 + t;
    const d3 v = mtv(Qm, ld3(rhs, a.ldb, m));
    st3(bnew, a.ldb, m, v);
    add(v, m, a.ldb);
  }
  if (blockIdx.x == sw_tile0[T.sw] && t < 2) {
    const int o = 3 * a.ldb + 6 * T.sw + 3 * t;
    const d3 v = mtv(Qm, ga3(rhs + o));
    sa3(bnew + o, v);
    add(v, o, 1);
  }
  const int lane = t & 63, w = t >> 6;
#pragma unroll
  for (int i = 0; i <= kHistMax; i++) {
    const double r = wave_sum(acc[i]);
    if (lane == 0) sh[4 * i + w] = r;
  }
  __syncthreads();
  if (t <= kHistMax)
    dpart[(kHistMax + 1) * size_t(blockIdx.x) + t] =
        ((sh[4 * t] + sh[4 * t + 1]) + sh[4 * t + 2]) + sh[4 * t + 3];
}

// Per swimmer (one thread): g = sum of the tile shares (tile order), then
// (G + lambda I) c = g by Cholesky, lambda = 1e-14 trace(G) / count.
// A failed factorisation (or an empty history) leaves c = 0 (no guess).  g
// and bnew . bnew are kept for the Gram update after the solve.
__global__ void k_hist_solve(SimDev a, const int *sw_tile0, const int *sw_ntile,
                             const double *dpart, const double *G, double *Gw, const int *hst,
                             double *coef, double *gsave) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= a.Nsw) return;
  const int count = hst[0];
  constexpr int M = kHistMax;
  double g[M + 1];
  for (int i = 0; i <= M; i++) {
    double s = 0.0;
    for (int q = sw_tile0[n]; q < sw_tile0[n] + sw_ntile[n]; q++) s += dpart[(M + 1) * size_t(q) + i];
    g[i] = s;
    gsave[(M + 1) * size_t(n) + i] = s;
  }
  const double *Gn = G + size_t(n) * M * M;
  double *L = Gw + size_t(n) * M * M;
  double tr = 0.0;
  for (int i = 0; i < count; i++) tr += Gn[M * i + i];
  const double lam = count > 0 ? 1e-14 * tr / count : 0.0;
  bool ok = tr > 0.0;
  for (int j = 0; j < count && ok; j++) {
    double d = Gn[M * j + j] + lam;
    for (int k = 0; k < j; k++) d -= L[M * j + k] * L[M * j + k];
    if (!(d > 0.0)) { ok = false; break; }
    const double ljj = sqrt(d);
    L[M * j + j] = ljj;
    for (int i = j + 1; i < count; i++) {
      double e = Gn[M * i + j];
      for (int k = 0; k < j; k++) e -= L[M * i + k] * L[M * j + k];
      L[M * i + j] = e / ljj;
    }
  }
  double c[M];
  for (int i = 0; i < M; i++) c[i] = 0.0;
  if (ok) {
    for (int i = 0; i < count; i++) {  // L y = g
      double e = g[i];
      for (int k = 0; k < i; k++) e -= L[M * i + k] * c[k];
      c[i] = e / L[M * i + i];
    }
    for (int i = count - 1; i >= 0; i--) {  // L^T c = y
      double e = c[i];
      for (int k = i + 1; k < count; k++) e -= L[M * k + i] * c[k];
      c[i] = e / L[M * i + i];
    }
  }
  for (int i = 0; i < M; i++) coef[M * size_t(n) + i] = ok ? c[i] : 0.0;
}

// Per tile: x = R sum_i c_i Hx[i] (the guess, rotated back to the lab frame);
// with an empty history x stays the previous solution.
__global__ __launch_bounds__(256) void k_hist_combine(SimDev a, const BTile *tiles,
                                                      const int *sw_tile0, const double *Hx,
                                                      size_t ldh, const int *hst,
                                                      const double *coef, double *x) {
  const int count = hst[0];
  if (count == 0) return;
  const BTile T = tiles[blockIdx.x];
  const int t = threadIdx.x;
  const m3 Qm = rotm(ga4(a.body + size_t(T.sw) * BREC + BQ));
  const double *c = coef + size_t(kHistMax) * T.sw;
  auto comb = [&](int o, int ld) -> d3 {
    d3 v = {0.0, 0.0, 0.0};
    for (int i = 0; i < count; i++) {
      const double *h = Hx + ldh * i + o;
      v = v + c[i] * d3{h[0], h[ld], h[2 * ld]};
    }
    return mv(Qm, v);
  };
  if (t < T.n) st3(x, a.ldb, T.m0 + t, comb(T.m0 + t, a.ldb));
  if (blockIdx.x == sw_tile0[T.sw] && t < 2) {
    const int o = 3 * a.ldb + 6 * T.sw + 3 * t;
    sa3(x + o, comb(o, 1));
  }
}

// After the solve: slot `head` <- (bnew, R^T x), and the swimmer's Gram
// matrix row and column of that slot from the saved dots.  hst = (count,
// head) before the push; k_hist_advance moves them on afterwards.
__global__ __launch_bounds__(256) void k_hist_push(SimDev a, const BTile *tiles,
                                                   const int *sw_tile0, const double *bnew,
                                                   const double *x, double *Hb, double *Hx,
                                                   size_t ldh, const int *hst, double *G,
                                                   const double *gsave, const int *skip) {
  if (skip && *skip) return;
  const int count_before = hst[0], head = hst[1];
  const BTile T = tiles[blockIdx.x];
  const int t = threadIdx.x;
  const m3 Qm = rotm(ga4(a.body + size_t(T.sw) * BREC + BQ));
  double *hb = Hb + ldh * head, *hx = Hx + ldh * head;
  if (t < T.n) {
    const int m = T.m0 + t;
    st3(hb, a.ldb, m, ld3(bnew, a.ldb, m));
    st3(hx, a.ldb, m, mtv(Qm, ld3(x, a.ldb, m)));
  }
  if (blockIdx.x == sw_tile0[T.sw]) {
    if (t < 2) {
      const int o = 3 * a.ldb + 6 * T.sw + 3 * t;
      sa3(hb + o, ga3(bnew + o));
      sa3(hx + o, mtv(Qm, ga3(x + o)));
    }
    if (t == 0) {
      constexpr int M = kHistMax;
      double *Gn = G + size_t(T.sw) * M * M;
      const double *g = gsave + (M + 1) * size_t(T.sw);
      for (int i = 0; i < count_before; i++)
        if (i != head) Gn[M * head + i] = Gn[M * i + head] = g[i];
      Gn[M * head + head] = g[M];
    }
  }
}

// the history's (count, head) after a push into a history of length m
__global__ void k_hist_advance(int *hst, int m, const int *skip) {
  if (skip && *skip) return;
  hst[0] = min(hst[0] + 1, m);
  hst[1] = (hst[1] + 1) % m;
}

// copy U of every swimmer into column `col` of N_ref
__global__ void k_store_nref_col(SimDev a, const double *x, int col) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= a.Nsw) return;
  double *b = a.body + size_t(n) * BREC;
  for (int r = 0; r < 6; r++) b[BN + 6 * r + col] = x[3 * a.ldb + 6 * n + r];
}

__global__ void k_unit_rhs(SimDev a, double *rhs, int slot) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.npad) return;
  double v = 0.0;
  if (i >= 3 * a.ldb && (i - 3 * a.ldb) % 6 == slot) v = -1.0;
  rhs[i] = v;
}

// -------------------------------------------------------------------------
// residual (mobility_solver.cpp:1404-1504, 1641-1689)

__global__ void k_errors(SimDev a, double *err, const double *vbody, const int *skip) {
  if (skip && *skip) return;
  const int nt = a.step_nt[0];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const double c1 = -4.0 / 3.0, c2 = 1.0 / 3.0, c3 = -2.0 / 3.0;
  const bool euler = nt < a.neuler;
  const double DT = a.dt;
  bool big = false, bad = false;
  if (i < a.Ns) {
    const int f = i / a.S, j = i % a.S;
    const int sw = a.fil_sw[f];
    const int id = 6 * a.S * f + 6 * sw + 3 * j;
    const int ld = a.lds;
    const d3 v = {a.Vs[i], a.Vs[ld + i], a.Vs[2 * ld + i]};
    const d3 om = {a.Vs[3 * ld + i], a.Vs[4 * ld + i], a.Vs[5 * ld + i]};
    const d3 x = ld3(a.X, ld, i), xm1 = ld3(a.Xm1, ld, i), xm2 = ld3(a.Xm2, ld, i);
    const d3 u = ld3(a.U, ld, i), um1 = ld3(a.Um1, ld, i);
    d3 e;
    if (euler) e = x - xm1 - DT * v;
    else e = x + c1 * xm1 + c2 * xm2 + c3 * DT * v;
    sa3(err + id, e);
    big = fabs(e.x) > a.DL * a.tol || fabs(e.y) > a.DL * a.tol || fabs(e.z) > a.DL * a.tol;
    bad = !(isfinite(e.x) && isfinite(e.y) && isfinite(e.z));
    const d3 W = dexpinv(u, om);
    d3 l;
    if (euler) l = u - DT * W;
    else l = u - c2 * um1 + c3 * DT * W;
    sa3(err + id + 3 * a.S, l);
    big = big || fabs(l.x) > a.tol || fabs(l.y) > a.tol || fabs(l.z) > a.tol;
    bad = bad || !(isfinite(l.x) && isfinite(l.y) && isfinite(l.z));
  } else if (i < a.Ns + a.Nsw) {
    const int n = i - a.Ns;
    const double *b = a.body + size_t(n) * BREC;
    const int id = 6 * a.S * (a.sw_fil0[n] + a.sw_nfil[n]) + 6 * n;
    const double *U = vbody + 6 * n;
    const d3 x = ga3(b + BX), xm1 = ga3(b + BXM1), xm2 = ga3(b + BXM2);
    const d3 u = ga3(b + BU), um1 = ga3(b + BUM1);
    const d3 w = dexpinv(u, rowvec(U + 3));
    d3 e, l;
    if (euler) {
      e = x - xm1 - DT * rowvec(U);
      l = u - DT * w;
    } else {
      e = x + c1 * xm1 + c2 * xm2 + c3 * DT * rowvec(U);
      l = u - c2 * um1 + c3 * DT * w;
    }
    sa3(err + id, e);
    sa3(err + id + 3, l);
    big = fabs(e.x) > a.tol * a.DL || fabs(e.y) > a.tol * a.DL || fabs(e.z) > a.tol * a.DL ||
          fabs(l.x) > a.tol || fabs(l.y) > a.tol || fabs(l.z) > a.tol;
    bad = !(isfinite(e.x) && isfinite(e.y) && isfinite(e.z) && isfinite(l.x) &&
            isfinite(l.y) && isfinite(l.z));
  }
  if (big) atomicOr(&a.flags[F_BIG], 1);
  if (bad) atomicOr(&a.flags[F_NAN], 1);
}

// -------------------------------------------------------------------------
// approximate Jacobian (filament.cpp:8-55, 1222-1539)

__device__ d3 psi_col(dq q, int c, double &s0) {
  // column c of psi_mat(q) (quaternion.cpp:315-329) as (s0, vec)
  const d3 v = {q.x, q.y, q.z};
  const d3 e = {double(c == 0), double(c == 1), double(c == 2)};
  s0 = -0.5 * dot3(v, e);
  const d3 r = mv(rcross(v), e);
  return 0.5 * (r + q.s * e);
}

// filament.cpp:8-55
__device__ m3 bfm_lie_derivative(const SimDev &a, dq q1, dq q2, bool upper) {
  const dq qm = qconj(qmid(q1, q2));
  const double s = 1.0 / a.DL;
  const dq dqs = {(q2.s - q1.s) * s, (q2.x - q1.x) * s, (q2.y - q1.y) * s, (q2.z - q1.z) * s};
  const dq qq = upper ? q2 : q1;
  const double s1 = upper ? 1.0 / a.DL : -1.0 / a.DL;
  const double s2 = upper ? 0.5 * a.DL : -0.5 * a.DL;
  m3 out;
  for (int c = 0; c < 3; c++) {
    double p0;
    const d3 pv = psi_col(qq, c, p0);
    const dq A1 = {s1 * p0, s1 * pv.x, s1 * pv.y, s1 * pv.z};
    const dq A2 = {s2 * A1.s, s2 * A1.x, s2 * A1.y, s2 * A1.z};
    const dq r1 = qmul(qm, A1), r2 = qmul(dqs, A2);
    out.a[0 * 3 + c] = 2.0 * (r1.x + r2.x) * a.kt;
    out.a[1 * 3 + c] = 2.0 * (r1.y + r2.y) * a.kb;
    out.a[2 * 3 + c] = 2.0 * (r1.z + r2.z) * a.kb;
  }
  return out;
}

__device__ m3 elastic_rc(d3 K, dq q) {
  return madd(mscale(K.x, rcross(qtangent(q))),
              madd(mscale(K.y, rcross(qnormal(q))), mscale(K.z, rcross(qbinormal(q)))));
}

__device__ void put3(double *J, int n, int r0, int c0, const m3 &B) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) J[size_t(r0 + i) * n + c0 + j] = B.a[3 * i + j];
}
__device__ m3 get3(const double *J, int n, int r0, int c0) {
  m3 B;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) B.a[3 * i + j] = J[size_t(r0 + i) * n + c0 + j];
  return B;
}

// Builds the 6S x 6S approximate Jacobian of filament blockIdx.x into Jw.
__global__ __launch_bounds__(64) void k_fil_jacobian(SimDev a, int nt, double t) {
  const int f = blockIdx.x;
  const int S = a.S, n = 6 * S;
  double *J = a.Jw + size_t(f) * n * n;
  for (int e = threadIdx.x; e < n * n; e += 64) J[e] = 0.0;
  __syncthreads();
  const int k = threadIdx.x;
  const bool euler = nt < a.neuler;
  const double mu = a.mu, as = a.as, DL = a.DL, dt = a.dt;
  const int g0 = f * S;
  if (k < S) {
    const int g = g0 + k;
    // positions <- lambda (filament.cpp:1414-1421)
    const double fac1 = euler ? -dt / (6.0 * kPi * mu * as) : -dt / (9.0 * kPi * mu * as);
    for (int c = 0; c < 3; c++) {
      const int i = 3 * k + c;
      J[size_t(i) * n + i] = fac1;
      if (i < 3 * S - 3) J[size_t(i) * n + i + 3] = -fac1;
    }
    // positions <- Lie algebra (robot arm, filament.cpp:1439-1455)
    for (int j = 1; j <= k; j++) {
      const m3 tm = mscale(DL * (j == k ? 0.5 : 1.0), rcross(seg_t(a, g0 + j)));
      put3(J, n, 3 * k, 3 * (S + j), tm);
    }
    // tip position <- tip Lie element through the end force (filament.cpp:1466-1470)
    if (a.efm != 0.0 && k == S - 1) {
      const m3 e = mscale(fac1 * a.efm, rcross(seg_t(a, g)));
      put3(J, n, 3 * k, 3 * (2 * S - 1), msub(get3(J, n, 3 * k, 3 * (2 * S - 1)), e));
    }
    // Lie <- lambda (filament.cpp:1463-1486)
    const double fac2 = euler ? DL * dt / (16.0 * kPi * mu * as * as * as)
                              : DL * dt / (24.0 * kPi * mu * as * as * as);
    {
      const m3 um = rcross(ld3(a.U, a.lds, g));
      const m3 tm = mscale(-1.0, rcross(seg_t(a, g)));
      const m3 blk = mscale(fac2, madd(tm, madd(mscale(0.5, mm(um, tm)),
                                                mscale(1.0 / 12.0, mm(mm(um, um), tm)))));
      if (k <= S - 2) put3(J, n, 3 * (S + k), 3 * (k + 1), blk);
      if (k >= 1) put3(J, n, 3 * (S + k), 3 * k, blk);
    }
    // Lie <- Lie (filament.cpp:1222-1404)
    const double tsf = euler ? dt : 2.0 * dt / 3.0;
    const double Tfac = 1.0 / (8.0 * kPi * mu * as * as * as);
    const d3 st = strain_twist(a, f, t);
    const d3 uk = ld3(a.U, a.lds, g);
    const m3 Lie = mscale(-1.0, rcross(uk));
    const d3 tau = {a.Fint[3 * a.lds + g], a.Fint[4 * a.lds + g], a.Fint[5 * a.lds + g]};
    const m3 Om = rcross(Tfac * tau);
    auto diag_block = [&](const m3 &dO) {
      const m3 t1 = madd(Om, mm(Lie, dO));
      const m3 t2 = madd(rcross(mv(Om, uk)), madd(mm(Lie, Om), mm(mm(Lie, Lie), dO)));
      return mscale(-tsf, madd(msub(dO, mscale(0.5, t1)), mscale(1.0 / 12.0, t2)));
    };
    auto off_block = [&](const m3 &dO) {
      return mscale(-tsf, madd(msub(dO, mscale(0.5, mm(Lie, dO))),
                               mscale(1.0 / 12.0, mm(mm(Lie, Lie), dO))));
    };
    const dq qk = seg_q(a, g);
    const d3 tk = qtangent(qk);
    const double *L = a.lam + size_t(f) * 9 * S;
    if (k == 0) {
      const dq q1 = seg_q(a, g + 1);
      const m3 Cp = mscale(-0.5 * DL, mm(rcross(rowvec(L)), rcross(tk)));
      const d3 Kph = bfm(a, qk, q1, st);
      const m3 R = rotm(qmid(qk, q1));
      m3 El = madd(mscale(0.5, elastic_rc(Kph, qk)), mm(R, bfm_lie_derivative(a, qk, q1, false)));
      put3(J, n, 3 * S, 3 * S, diag_block(mscale(Tfac, madd(El, Cp))));
      El = madd(mscale(0.5, elastic_rc(Kph, q1)), mm(R, bfm_lie_derivative(a, qk, q1, true)));
      put3(J, n, 3 * S, 3 * S + 3, off_block(mscale(Tfac, El)));
    } else if (k == S - 1) {
      const dq q0 = seg_q(a, g - 1);
      const m3 Cp = mscale(-0.5 * DL, mm(rcross(rowvec(L + 3 * (S - 2))), rcross(tk)));
      const d3 Kmh = bfm(a, q0, qk, st);
      const m3 R = rotm(qmid(q0, qk));
      m3 El = msub(mscale(-0.5, elastic_rc(Kmh, qk)), mm(R, bfm_lie_derivative(a, q0, qk, true)));
      put3(J, n, 3 * S + 3 * k, 3 * S + 3 * k, diag_block(mscale(Tfac, madd(El, Cp))));
      El = msub(mscale(-0.5, elastic_rc(Kmh, q0)), mm(R, bfm_lie_derivative(a, q0, qk, false)));
      put3(J, n, 3 * S + 3 * k, 3 * S + 3 * k - 3, off_block(mscale(Tfac, El)));
    } else {
      const dq qm = seg_q(a, g - 1), qp = seg_q(a, g + 1);
      const d3 Kph = bfm(a, qk, qp, st), Kmh = bfm(a, qm, qk, st);
      const d3 lv = rowvec(L + 3 * (k - 1)) + rowvec(L + 3 * k);
      const m3 Cp = mscale(-0.5 * DL, mm(rcross(lv), rcross(tk)));
      const m3 Rp = rotm(qmid(qk, qp)), Rm = rotm(qmid(qm, qk));
      m3 El = madd(mscale(0.5, elastic_rc(Kph, qk)), mm(Rp, bfm_lie_derivative(a, qk, qp, false)));
      El = msub(El, mscale(0.5, elastic_rc(Kmh, qk)));
      El = msub(El, mm(Rm, bfm_lie_derivative(a, qm, qk, true)));
      put3(J, n, 3 * S + 3 * k, 3 * S + 3 * k, diag_block(mscale(Tfac, madd(Cp, El))));
      El = msub(mscale(-0.5, elastic_rc(Kmh, qm)), mm(Rm, bfm_lie_derivative(a, qm, qk, false)));
      put3(J, n, 3 * S + 3 * k, 3 * S + 3 * k - 3, off_block(mscale(Tfac, El)));
      El = madd(mscale(0.5, elastic_rc(Kph, qp)), mm(Rp, bfm_lie_derivative(a, qk, qp, true)));
      put3(J, n, 3 * S + 3 * k, 3 * S + 3 * k + 3, off_block(mscale(Tfac, El)));
    }
    for (int c = 0; c < 3; c++) J[size_t(3 * S + 3 * k + c) * n + 3 * S + 3 * k + c] += 1.0;
  }
  __syncthreads();
  // clamped first frame (filament.cpp:1496-1532)
  if (threadIdx.x == 0) {
    const double fac3 = euler ? -dt / (8.0 * kPi * mu * as * as * as)
                              : -dt / (12.0 * kPi * mu * as * as * as);
    const m3 um = rcross(ld3(a.U, a.lds, g0));
    const m3 blk = madd(meye(), madd(mscale(0.5, um), mscale(1.0 / 12.0, mm(um, um))));
    const m3 blkT = madd(meye(), madd(mscale(-0.5, um), mscale(1.0 / 12.0, mm(um, um))));
    const m3 e1 = get3(J, n, 3 * S, 3 * S), e2 = get3(J, n, 3 * S + 3, 3 * S);
    double *E = a.ecb + 18 * f;
    for (int q = 0; q < 9; q++) { E[q] = e1.a[q]; E[9 + q] = e2.a[q]; }
    put3(J, n, 3 * S, 3 * S, mscale(fac3, mm(blk, blkT)));
    put3(J, n, 3 * S + 3, 3 * S, mzero());
  }
}

// In-place Gauss-Jordan inverse with partial pivoting of each filament's
// matrix (Jw), written column-unpermuted to Jinv.  Workgroup per filament.
__global__ __launch_bounds__(256) void k_gj_invert(SimDev a) {
  const int f = blockIdx.x;
  const int n = 6 * a.S;
  double *A = a.Jw + size_t(f) * n * n;
  double *Out = a.Jinv + size_t(f) * n * n;
  __shared__ int piv[6 * kMaxSeg];
  __shared__ double colk[6 * kMaxSeg];
  __shared__ double rv[256];
  __shared__ int ri[256];
  const int t = threadIdx.x;
  for (int i = t; i < n; i += 256) piv[i] = i;
  __syncthreads();
  for (int k = 0; k < n; k++) {
    // pivot: first row of maximal |A(i,k)|, i >= k
    double best = -1.0;
    int bi = n;
    for (int i = k + t; i < n; i += 256) {
      const double v = fabs(A[size_t(i) * n + k]);
      if (v > best) { best = v; bi = i; }
    }
    rv[t] = best; ri[t] = bi;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (t < s) {
        const double v2 = rv[t + s];
        const int i2 = ri[t + s];
        if (v2 > rv[t] || (v2 == rv[t] && i2 < ri[t])) { rv[t] = v2; ri[t] = i2; }
      }
      __syncthreads();
    }
    const int p = ri[0];
    const double pv = rv[0];
    __syncthreads();
    if (!(pv > 0.0)) {
      if (t == 0) atomicOr(&a.flags[F_SING], 1);
      return;
    }
    if (p != k) {
      for (int j = t; j < n; j += 256) {
        const double x = A[size_t(k) * n + j];
        A[size_t(k) * n + j] = A[size_t(p) * n + j];
        A[size_t(p) * n + j] = x;
      }
      if (t == 0) { const int x = piv[k]; piv[k] = piv[p]; piv[p] = x; }
    }
    __syncthreads();
    const double inv = 1.0 / A[size_t(k) * n + k];
    for (int i = t; i < n; i += 256) colk[i] = A[size_t(i) * n + k];
    __syncthreads();
    for (int j = t; j < n; j += 256)
      A[size_t(k) * n + j] = (j == k ? 1.0 : A[size_t(k) * n + j]) * inv;
    __syncthreads();
    for (int e = t; e < n * n; e += 256) {
      const int i = e / n, j = e % n;
      if (i == k) continue;
      const double fi = colk[i];
      const double base = (j == k) ? 0.0 : A[e];
      A[e] = base - fi * A[size_t(k) * n + j];
    }
    __syncthreads();
  }
  for (int e = t; e < n * n; e += 256) {
    const int i = e / n, j = e % n;
    Out[size_t(i) * n + piv[j]] = A[e];
  }
}

// The same Gauss-Jordan inverse with the matrix held in LDS (n*n doubles,
// 115 KB at NSEG = 20; dynamic shared memory), 1024 threads: thread
// (column j = t % 128, row group t / 128) updates rows j-group, +8, ...
// Used when the device's per-workgroup LDS holds the matrix.
constexpr int kGjThreads = 1024;

__global__ __launch_bounds__(kGjThreads) void k_gj_invert_lds(SimDev a) {
  extern __shared__ double lds_gj[];
  __shared__ int sp;
  __shared__ double spv, spk, skk;  // |pivot|, pivot A(p,k), A(k,k) before the swap
  const int f = blockIdx.x;
  const int n = 6 * a.S;
  const int t = threadIdx.x;
  double *A = lds_gj;
  double *colk = A + size_t(n) * n;
  int *piv = reinterpret_cast<int *>(colk + n);
  const double *G = a.Jw + size_t(f) * n * n;
  for (int e = t; e < n * n; e += kGjThreads) A[e] = G[e];
  for (int i = t; i < n; i += kGjThreads) piv[i] = i;
  const int jc = t & 127, rg = t >> 7;
  __syncthreads();
  for (int k = 0; k < n; k++) {
    // pivot: first row of maximal |A(i,k)|, i >= k (wave 0)
    if (t < 64) {
      double best = -1.0;
      int bi = n;
      for (int i = k + t; i < n; i += 64) {
        const double v = fabs(A[size_t(i) * n + k]);
        if (v > best) { best = v; bi = i; }
      }
      for (int o = 32; o > 0; o >>= 1) {
        const double v2 = __shfl_xor(best, o, 64);
        const int i2 = __shfl_xor(bi, o, 64);
        if (v2 > best || (v2 == best && i2 < bi)) { best = v2; bi = i2; }
      }
      if (t == 0) {
        sp = bi;
        spv = best;
        if (bi < n) {
          spk = A[size_t(bi) * n + k];
          skk = A[size_t(k) * n + k];
        }
      }
    }
    __syncthreads();
    const int p = sp;
    if (!(spv > 0.0)) {
      if (t == 0) atomicOr(&a.flags[F_SING], 1);
      return;
    }
    // swap rows k and p, normalise the new row k and keep column k (as it is
    // after the swap) in one pass: only rows k and p change here, and their
    // column-k entries come from the pivot search
    const double inv = 1.0 / spk;
    if (t < n) {
      const double akj = A[size_t(k) * n + t], apj = A[size_t(p) * n + t];
      A[size_t(k) * n + t] = (t == k ? 1.0 : apj) * inv;
      if (p != k) A[size_t(p) * n + t] = akj;
    } else if (t >= 512 && t - 512 < n) {
      const int i = t - 512;
      colk[i] = i == k ? spk : (i == p ? skk : A[size_t(i) * n + k]);
    }
    if (t == 0 && p != k) { const int x = piv[k]; piv[k] = piv[p]; piv[p] = x; }
    __syncthreads();
    // columns jc, jc + 128, ... (n exceeds 128 from NSEG = 22 on)
    for (int j = jc; j < n; j += 128) {
      const double akj = A[size_t(k) * n + j];
      for (int i = rg; i < n; i += kGjThreads / 128) {
        if (i == k) continue;
        const double base = (j == k) ? 0.0 : A[size_t(i) * n + j];
        A[size_t(i) * n + j] = base - colk[i] * akj;
      }
    }
    __syncthreads();
  }
  double *Out = a.Jinv + size_t(f) * n * n;
  for (int e = t; e < n * n; e += kGjThreads) {
    const int i = e / n, j = e % n;
    Out[size_t(i) * n + piv[j]] = A[e];
  }
}

inline size_t gj_lds_bytes(int S) {
  const size_t n = 6 * size_t(S);
  return n * n * sizeof(double) + n * sizeof(double) + n * sizeof(int);
}

// The same Gauss-Jordan inverse with the matrix in registers, for n <= 128
// (NSEG <= 21).  256 threads, one wave per SIMD: thread t holds rows
// t / 8 + 32 rho (rho < 4) and columns t % 8 + 8u (u < 16), so each step's
// pivot row, read once from LDS, serves four rows, and the per-step
// bookkeeping runs on four waves only.  Rows are not swapped: step k takes
// as pivot an unused row of largest |A(i,k)|, so row p_k of the in-place
// result is row k of the inverse of the row-permuted matrix, and
// Out(k, p_j) = A(p_k, j).
//
// The pivot is the largest key over the candidate rows: |A(i,k)| (clamped
// to the largest finite double) with the low 7 mantissa bits replaced by
// 127 - i, compared as a double (a non-negative double orders as its bit
// pattern), so the lowest row wins among magnitudes that agree to 2^-45;
// each wave takes its maximum by DPP and publishes it with the inverse of
// its candidate's value, computed before the barrier.
//
// One step, k: the elimination of step k - 1 is applied first to the slot
// u = k / 8 that holds column k (the outer loop is unrolled, so that slot is
// a register), column k and the wave keys are published (barrier), the
// rest of step k - 1's elimination runs while the keys are read, the pivot
// row normalises and publishes itself (barrier).  Every entry sees the same
// operations as in the in-LDS kernel; two buffers let step k + 1 publish
// while step k is still being read.  (This takes about 1.3 us per step at
// n = 120: the chain of dependent LDS, DPP and FP64 operations between the
// barriers, not the arithmetic, sets the time.)
constexpr int kGjRegMax = 128;
constexpr int kGjRegCols = kGjRegMax / 8;
constexpr int kGjRegRows = 4;

struct GjRegShared {
  double scol[2][kGjRegMax], srow[2][kGjRegMax];
  double skey[2][4], sinv[2][4];  // per wave: best key, inverse of its pivot
  int prow[kGjRegMax];
};

struct GjRegState {
  double v[kGjRegRows][kGjRegCols];  // rows t / 8 + 32 rho, columns t % 8 + 8u
  bool used[kGjRegRows];
  int mystep[kGjRegRows];
  int plast;  // pivot row of the last step, whose elimination is pending
};

// The eliminations of step kp (pivot row st.plast, its normalised row and
// column kp in buffer kp & 1) on slots u in [U0, U1), skipping slot SKIP;
// column kp lies in a slot of [C0, C1), the only ones that need a select.
template <int U0, int U1, int SKIP, int C0, int C1>
__device__ __forceinline__ void gj_reg_eliminate(GjRegState &st, const double (&pr)[kGjRegCols],
                                                 const double (&fi)[kGjRegRows], int kp) {
  const int g = threadIdx.x & 7;
#pragma unroll
  for (int q = 0; q < kGjRegRows; q++) {
    const bool piv = (threadIdx.x >> 3) + 32 * q == st.plast;
#pragma unroll
    for (int u = U0; u < U1; u++) {
      if (u == SKIP) continue;
      if (u >= C0 && u < C1) {
        const bool colk = g + 8 * u == kp && !piv;
        st.v[q][u] = (colk ? 0.0 : st.v[q][u]) - fi[q] * pr[u];
      } else {
        st.v[q][u] = st.v[q][u] - fi[q] * pr[u];
      }
    }
  }
}

__device__ __forceinline__ void gj_reg_pending(const GjRegShared &sh, const GjRegState &st, int kp,
                                               double (&pr)[kGjRegCols], double (&fi)[kGjRegRows]) {
  const int t = threadIdx.x, g = t & 7, b = kp & 1;
#pragma unroll
  for (int u = 0; u < kGjRegCols; u++) pr[u] = sh.srow[b][g + 8 * u];
#pragma unroll
  for (int q = 0; q < kGjRegRows; q++) {
    const int r = (t >> 3) + 32 * q;
    fi[q] = r == st.plast ? 0.0 : sh.scol[b][r];
  }
}

// Steps k = 8 UK + gk, gk < 8, k < n; false if a pivot is zero or not finite.
template <int UK>
__device__ __forceinline__ bool gj_reg_block(GjRegShared &sh, GjRegState &st, int n) {
  const int t = threadIdx.x, g = t & 7, lane = t & 63, wave = t >> 6;
  for (int gk = 0; gk < 8; gk++) {
    const int k = 8 * UK + gk;
    if (k >= n) break;
    const int b = k & 1;
    double pr[kGjRegCols], fi[kGjRegRows];
    if (k > 0) {
      gj_reg_pending(sh, st, k - 1, pr, fi);
      gj_reg_eliminate<UK, UK + 1, -1, UK, UK + 1>(st, pr, fi, k - 1);
    }
    // owner lanes (g = gk): the best of their four rows as a key and the
    // inverse of its value (used only if that row wins)
    double key = 0.0, kinv = 0.0;
    if (g == gk) {
      double xb = 0.0;
#pragma unroll
      for (int q = 0; q < kGjRegRows; q++) {
        const int r = (t >> 3) + 32 * q;
        const double x = st.v[q][UK];
        sh.scol[b][r] = x;
        const double ax = fabs(x);
        const unsigned long long bits = __double_as_longlong(ax <= DBL_MAX ? ax : DBL_MAX);
        const double kq = __longlong_as_double((long long)((bits & ~0x7full) |
                                                           (unsigned long long)(127 - r)));
        if (!st.used[q] && kq > key) { key = kq; xb = x; }
      }
      kinv = 1.0 / xb;
    }
    // the wave's best: owner lanes gk + 8m pair up by a rotation of 8 within
    // each row of 16 lanes, then the four rows
    key = fmax(key, dpp_d<0x128>(key));  // row_ror:8
    key = fmax(fmax(lane_d(key, gk), lane_d(key, gk + 16)),
               fmax(lane_d(key, gk + 32), lane_d(key, gk + 48)));
    {
      const int rw = 127 - int(__double_as_longlong(key) & 0x7f);
      const double winv = lane_d(kinv, 8 * (rw & 7) + gk);
      if (lane == 0) { sh.skey[b][wave] = key; sh.sinv[b][wave] = winv; }
    }
    __syncthreads();
    double kw[4], iw[4];
#pragma unroll
    for (int w = 0; w < 4; w++) { kw[w] = sh.skey[b][w]; iw[w] = sh.sinv[b][w]; }
    // column k - 1 is in slot UK, or UK - 1 when gk = 0
    if (k > 0) gj_reg_eliminate<0, kGjRegCols, UK, UK - 1, UK>(st, pr, fi, k - 1);
    double inv;
    {
      const bool a01 = kw[0] > kw[1], a23 = kw[2] > kw[3];
      const double k01 = a01 ? kw[0] : kw[1], k23 = a23 ? kw[2] : kw[3];
      const double i01 = a01 ? iw[0] : iw[1], i23 = a23 ? iw[2] : iw[3];
      key = k01 > k23 ? k01 : k23;
      inv = k01 > k23 ? i01 : i23;
    }
    // a zero, infinite or NaN pivot: singular
    if (!(key > 0.0) || !(fabs(inv) <= DBL_MAX) || inv == 0.0) return false;
    const int p = 127 - int(__double_as_longlong(key) & 0x7f);
#pragma unroll
    for (int q = 0; q < kGjRegRows; q++) {
      if ((t >> 3) + 32 * q == p) {
#pragma unroll
        for (int u = 0; u < kGjRegCols; u++) {
          st.v[q][u] = ((u == UK && g == gk) ? 1.0 : st.v[q][u]) * inv;
          sh.srow[b][g + 8 * u] = st.v[q][u];
        }
        st.used[q] = true;
        st.mystep[q] = k;
        if (g == 0) sh.prow[k] = p;
      }
    }
    st.plast = p;
    __syncthreads();
  }
  return true;
}

template <int... UK>
__device__ __forceinline__ bool gj_reg_all(GjRegShared &sh, GjRegState &st, int n,
                                           std::integer_sequence<int, UK...>) {
  // blocks in order; a block past n does nothing
  return ((8 * UK >= n || gj_reg_block<UK>(sh, st, n)) && ...);
}

__global__ __launch_bounds__(256) void k_gj_invert_reg(SimDev a, const int *redo) {
  __shared__ GjRegShared sh;
  const int f = blockIdx.x;
  if (redo && !redo[f]) return;
  const int n = 6 * a.S;
  const int t = threadIdx.x, g = t & 7;
  const double *G = a.Jw + size_t(f) * n * n;
  GjRegState st;
  st.plast = -1;
#pragma unroll
  for (int q = 0; q < kGjRegRows; q++) {
    const int r = (t >> 3) + 32 * q;
    st.used[q] = r >= n;
    st.mystep[q] = 0;
#pragma unroll
    for (int u = 0; u < kGjRegCols; u++) {
      const int c = g + 8 * u;
      st.v[q][u] = (r < n && c < n) ? G[size_t(r) * n + c] : 0.0;
    }
  }
  if (!gj_reg_all(sh, st, n, std::make_integer_sequence<int, kGjRegCols>{})) {
    if (t == 0) atomicOr(&a.flags[F_SING], 1);
    return;
  }
  {
    // the last step's elimination
    double pr[kGjRegCols], fi[kGjRegRows];
    gj_reg_pending(sh, st, n - 1, pr, fi);
    gj_reg_eliminate<0, kGjRegCols, -1, 0, kGjRegCols>(st, pr, fi, n - 1);
  }
#pragma unroll
  for (int q = 0; q < kGjRegRows; q++) {
    const int r = (t >> 3) + 32 * q;
    if (r >= n) continue;
    double *Out = a.Jinv + size_t(f) * n * n + size_t(st.mystep[q]) * n;
#pragma unroll
    for (int u = 0; u < kGjRegCols; u++) {
      const int c = g + 8 * u;
      if (c < n) Out[sh.prow[c]] = st.v[q][u];
    }
  }
}

// Gauss-Jordan without pivoting, same register layout as k_gj_invert_reg,
// one barrier per step.  The filament Jacobians are inverted this way first:
// on them the pivots stay within four orders of magnitude of their column's
// largest entry, and the inverse's residual is within a factor of a few of
// the pivoted one's.  Each inverse is checked on a fixed vector x with
// entries in [1, 2): y = J x is formed from the loaded matrix, and
// gj_redo[f] = 1 unless max |J^-1 y - x| <= 2e-6 (a NaN fails); then
// k_gj_invert_reg (partial pivoting) redoes the filaments that failed.
// Rounding alone leaves up to about cond(J) x 1e-16 there (8e-10 for the
// pivoted and 3e-9 for the pivot-free inverses of the 962-particle
// swimmer's filaments, cond ~3e6), so the check passes rounding and catches
// an elimination that lost the matrix.
// Step k finds in LDS the normalised pivot row k and column k of the
// partially eliminated matrix, published by the owners at the end of step
// k - 1; every thread eliminates its entries (the pivot row is already its
// normalised copy), then the owners of row k + 1 normalise it and the owners
// of column k + 1 publish it for the next step.
struct GjNpShared {
  double srow[2][kGjRegMax], scol[2][kGjRegMax];
  double y[kGjRegMax];  // J x of the check
  int fail;
};

__device__ __forceinline__ double gj_check_x(int c) { return 1.0 + double((c * 37) % 64) / 64.0; }

// Sum over the 8 lanes holding one row (lanes 8m .. 8m + 7), in each of them.
__device__ __forceinline__ double row8_sum(double v) {
  v += dpp_d<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_d<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_d<0x141>(v);  // row_half_mirror
  return v;
}

// The owners of column kp (slot u of each held row) and of row kp publish
// step kp's column and normalised row into buffer kp & 1.
template <int SLOT>
__device__ __forceinline__ void gj_np_publish(GjNpShared &sh, double (&v)[kGjRegRows][kGjRegCols],
                                              int kp) {
  const int t = threadIdx.x, g = t & 7, b = kp & 1;
  if (g == (kp & 7)) {
#pragma unroll
    for (int q = 0; q < kGjRegRows; q++) sh.scol[b][(t >> 3) + 32 * q] = v[q][SLOT];
  }
  if ((t >> 6) == ((kp & 31) >> 3)) {  // the wave holding row kp
#pragma unroll
    for (int q = 0; q < kGjRegRows; q++) {
      if (q != (kp >> 5)) continue;
      // the pivot from the lane of row kp that holds column kp
      const double inv = 1.0 / lane_d(v[q][SLOT], 9 * (kp & 7));
      if ((t >> 3) == (kp & 31)) {
#pragma unroll
        for (int u = 0; u < kGjRegCols; u++) {
          v[q][u] = ((u == SLOT && g == (kp & 7)) ? 1.0 : v[q][u]) * inv;
          sh.srow[b][g + 8 * u] = v[q][u];
        }
      }
    }
  }
}

template <int UK>
__device__ __forceinline__ void gj_np_block(GjNpShared &sh, double (&v)[kGjRegRows][kGjRegCols],
                                            int n) {
  const int t = threadIdx.x, g = t & 7;
  for (int gk = 0; gk < 8; gk++) {
    const int k = 8 * UK + gk;
    if (k >= n) break;
    const int b = k & 1;
    double pr[kGjRegCols], fi[kGjRegRows];
#pragma unroll
    for (int u = 0; u < kGjRegCols; u++) pr[u] = sh.srow[b][g + 8 * u];
#pragma unroll
    for (int q = 0; q < kGjRegRows; q++) {
      const int r = (t >> 3) + 32 * q;
      fi[q] = r == k ? 0.0 : sh.scol[b][r];
    }
#pragma unroll
    for (int q = 0; q < kGjRegRows; q++) {
      const bool piv = (t >> 3) + 32 * q == k;
#pragma unroll
      for (int u = 0; u < kGjRegCols; u++) {
        if (u == UK)
          v[q][u] = ((g == gk && !piv) ? 0.0 : v[q][u]) - fi[q] * pr[u];
        else
          v[q][u] = v[q][u] - fi[q] * pr[u];
      }
    }
    if (k + 1 < n) {
      if (gk < 7) gj_np_publish<UK>(sh, v, k + 1);
      else if (UK + 1 < kGjRegCols) gj_np_publish<(UK + 1 < kGjRegCols ? UK + 1 : UK)>(sh, v, k + 1);
    }
    __syncthreads();
  }
}

template <int... UK>
__device__ __forceinline__ void gj_np_all(GjNpShared &sh, double (&v)[kGjRegRows][kGjRegCols],
                                          int n, std::integer_sequence<int, UK...>) {
  ((8 * UK < n ? gj_np_block<UK>(sh, v, n) : void()), ...);
}

__global__ __launch_bounds__(256) void k_gj_nopiv_reg(SimDev a) {
  __shared__ GjNpShared sh;
  const int f = blockIdx.x;
  const int n = 6 * a.S;
  const int t = threadIdx.x, g = t & 7;
  const double *G = a.Jw + size_t(f) * n * n;
  double v[kGjRegRows][kGjRegCols];
#pragma unroll
  for (int q = 0; q < kGjRegRows; q++) {
    const int r = (t >> 3) + 32 * q;
#pragma unroll
    for (int u = 0; u < kGjRegCols; u++) {
      const int c = g + 8 * u;
      v[q][u] = (r < n && c < n) ? G[size_t(r) * n + c] : 0.0;
    }
  }
  // the check's y = J x (rows past n hold zeros)
#pragma unroll
  for (int q = 0; q < kGjRegRows; q++) {
    double sq = 0.0;
#pragma unroll
    for (int u = 0; u < kGjRegCols; u++) sq += v[q][u] * gj_check_x(g + 8 * u);
    sq = row8_sum(sq);
    if (g == 0) sh.y[(t >> 3) + 32 * q] = sq;
  }
  if (t == 0) sh.fail = 0;
  gj_np_publish<0>(sh, v, 0);
  __syncthreads();
  gj_np_all(sh, v, n, std::make_integer_sequence<int, kGjRegCols>{});
  // the check: J^-1 y against x, row by row
  bool fail = false;
#pragma unroll
  for (int q = 0; q < kGjRegRows; q++) {
    double sq = 0.0;
#pragma unroll
    for (int u = 0; u < kGjRegCols; u++) sq += v[q][u] * sh.y[g + 8 * u];
    sq = row8_sum(sq);
    const int r = (t >> 3) + 32 * q;
    if (r < n && !(fabs(sq - gj_check_x(r)) <= 2e-6)) fail = true;
  }
  if (fail) sh.fail = 1;
  __syncthreads();
  if (t == 0) a.gj_redo[f] = sh.fail;
  double *Out = a.Jinv + size_t(f) * n * n;
#pragma unroll
  for (int q = 0; q < kGjRegRows; q++) {
    const int r = (t >> 3) + 32 * q;
    if (r >= n) continue;
#pragma unroll
    for (int u = 0; u < kGjRegCols; u++) {
      const int c = g + 8 * u;
      if (c < n) Out[size_t(r) * n + c] = v[q][u];
    }
  }
}

// Schur-complement blocks of one filament (swimmer.cpp:526-680):
// Bblk = B J^-1 (6 x 6S) and BA = Bblk A (6 x 6).
__global__ __launch_bounds__(256) void k_schur_fil(SimDev a, int nt) {
  const int f = blockIdx.x;
  const int S = a.S, n = 6 * S;
  __shared__ double B[6 * 6 * kMaxSeg];
  const int t = threadIdx.x;
  const double *b = a.body + size_t(a.fil_sw[f]) * BREC;
  const bool euler = nt < a.neuler;
  const double dtf = euler ? -a.dt : -2.0 * a.dt / 3.0;
  m3 Qi;
  for (int q = 0; q < 9; q++) Qi.a[q] = b[BQI + q];
  const d3 ub = ga3(b + BU);
  const m3 blc = mscale(-1.0, rcross(ub));
  const m3 Dinv = madd(meye(), madd(mscale(1.0 / 12.0, mm(blc, blc)), mscale(-0.5, blc)));
  const m3 DinvT = madd(meye(), madd(mscale(1.0 / 12.0, mm(blc, blc)), mscale(0.5, blc)));
  const d3 d = mv(Qi, rowvec(a.fil_ref + 3 * f));
  for (int e = t; e < 6 * n; e += 256) B[e] = 0.0;
  __syncthreads();
  if (t == 0) {
    double bm[36];
    rotate66(b + BN, Qi, bm);
    // raw B: cols 0..2 = [-I ; rcross(d)], cols 3S..3S+2 = [0 ; -DinvT]
    double raw[6][6];
    for (int r = 0; r < 6; r++)
      for (int c = 0; c < 6; c++) raw[r][c] = 0.0;
    const m3 rd = rcross(d);
    for (int i = 0; i < 3; i++) {
      raw[i][i] = -1.0;
      for (int j = 0; j < 3; j++) {
        raw[3 + i][j] = rd.a[3 * i + j];
        raw[3 + i][3 + j] = -DinvT.a[3 * i + j];
      }
    }
    for (int r = 0; r < 6; r++)
      for (int c = 0; c < 6; c++) {
        double s = 0;
        for (int k = 0; k < 6; k++) s += bm[6 * r + k] * raw[k][c];
        B[r * n + (c < 3 ? c : 3 * S + c - 3)] = s;
      }
    if (a.efm != 0.0) {
      // the end force's couplings to the Lie elements (swimmer.cpp:557-571):
      // tip column [Ef ; rcross(DL/2 t - (x_tip - x)) Ef], columns of
      // segments 1..S-2 [0 ; Ef rcross(DL t_m)], Ef = rcross(efm t_tip);
      // then premultiplied by the rotated body mobility like the rest of B
      const int gt = f * S + S - 1;
      const d3 tt = seg_t(a, gt);
      const m3 Ef = rcross(a.efm * tt);
      const d3 arm = (0.5 * a.DL) * tt - (ld3(a.X, a.lds, gt) - ga3(b + BX));
      for (int m = 1; m < S; m++) {
        m3 top = mzero(), bot;
        if (m == S - 1) {
          top = Ef;
          bot = mm(rcross(arm), Ef);
        } else {
          bot = mm(Ef, rcross(a.DL * seg_t(a, f * S + m)));
        }
        const int c0 = 3 * (S + m);
        for (int r = 0; r < 6; r++)
          for (int j = 0; j < 3; j++) {
            double s = 0;
            for (int k = 0; k < 3; k++)
              s += bm[6 * r + k] * top.a[3 * k + j] + bm[6 * r + 3 + k] * bot.a[3 * k + j];
            B[r * n + c0 + j] = s;
          }
      }
    }
    // sphere-segment RPY corrections (swimmer.cpp:580-650)
    const double R = a.sw_R[a.fil_sw[f]], mu = a.mu, as = a.as;
    const d3 bx = ga3(b + BX);
    auto add3 = [&](int r0, int c0, const m3 &M) {
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) B[(r0 + i) * n + c0 + j] += M.a[3 * i + j];
    };
    for (int s = 0; s < S; s++) {
      const int g = f * S + s;
      d3 rh = ld3(a.X, a.lds, g) - bx;
      const double r = sqrt(dot3(rh, rh));
      rh = (1.0 / r) * rh;
      m3 rr;
      for (int i = 0; i < 3; i++) {
        const double ri = i == 0 ? rh.x : (i == 1 ? rh.y : rh.z);
        rr.a[3 * i] = ri * rh.x; rr.a[3 * i + 1] = ri * rh.y; rr.a[3 * i + 2] = ri * rh.z;
      }
      const double sig = R * R + as * as;
      const m3 Mtt = mscale(1.0 / (8.0 * kPi * mu * r),
                            madd(mscale(1.0 + sig / (3.0 * r * r), meye()),
                                 mscale(1.0 - sig / (r * r), rr)));
      const m3 Mtr = mscale(1.0 / (8.0 * kPi * mu * r * r), rcross(rh));
      const m3 Mrr = mscale(1.0 / (16.0 * kPi * mu * r * r * r), msub(mscale(3.0, rr), meye()));
      const m3 tm = rcross(seg_t(a, g));
      if (s == 0) {
        add3(0, 0, Mtt);
        add3(3, 0, Mtr);
        add3(0, 3 * S, mm(Mtr, DinvT));
        add3(3, 3 * S, mm(Mrr, DinvT));
        add3(0, 3, msub(mscale(0.5 * a.DL, mm(Mtr, tm)), Mtt));
        add3(3, 3, msub(mscale(0.5 * a.DL, mm(Mrr, tm)), Mtr));
      } else if (s == S - 1) {
        if (a.efm != 0.0) {  // swimmer.cpp:630-633
          add3(0, 6 * S - 3, mscale(-a.efm, mm(Mtt, tm)));
          add3(3, 6 * S - 3, mscale(-a.efm, mm(Mtr, tm)));
        }
        add3(0, 3 * S - 3, madd(mscale(0.5 * a.DL, mm(Mtr, tm)), Mtt));
        add3(3, 3 * S - 3, madd(mscale(0.5 * a.DL, mm(Mrr, tm)), Mtr));
      } else {
        add3(0, 3 * s, madd(mscale(0.5 * a.DL, mm(Mtr, tm)), Mtt));
        add3(3, 3 * s, madd(mscale(0.5 * a.DL, mm(Mrr, tm)), Mtr));
        add3(0, 3 * (s + 1), msub(mscale(0.5 * a.DL, mm(Mtr, tm)), Mtt));
        add3(3, 3 * (s + 1), msub(mscale(0.5 * a.DL, mm(Mrr, tm)), Mtr));
      }
    }
  }
  __syncthreads();
  // rows 3..5 <- Dinv * rows 3..5 for every column block, then *dtf
  for (int cb = t; cb < 2 * S; cb += 256) {
    const int c0 = 3 * cb;
    m3 blk;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) blk.a[3 * i + j] = B[(3 + i) * n + c0 + j];
    const m3 r = mm(Dinv, blk);
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) B[(3 + i) * n + c0 + j] = r.a[3 * i + j];
  }
  __syncthreads();
  for (int e = t; e < 6 * n; e += 256) B[e] *= dtf;
  __syncthreads();
  // Bblk = B J^-1
  const double *Ji = a.Jinv + size_t(f) * n * n;
  double *Bb = a.Bblk + size_t(f) * 6 * n;
  for (int e = t; e < 6 * n; e += 256) {
    const int r = e / n, c = e % n;
    double s = 0;
    for (int k = 0; k < n; k++) s += B[r * n + k] * Ji[size_t(k) * n + c];
    Bb[e] = s;
  }
  __syncthreads();
  // BA = Bblk A (A as in swimmer.cpp:535-541,673-680)
  if (t < 36) {
    const int r = t / 6, c = t % 6;
    const double *E = a.ecb + 18 * f;
    const double dn = sqrt(dot3(d, d));
    const d3 v = ((0.5 * a.DL + dn) / dn) * d;
    const m3 Dm = rcross(v);
    double s = 0;
    if (c < 3) {
      for (int k = 0; k < S; k++) s += Bb[r * n + 3 * k + c];
    } else {
      const int cc = c - 3;
      for (int k = 0; k < S; k++)
        for (int q = 0; q < 3; q++) s += Bb[r * n + 3 * k + q] * Dm.a[3 * q + cc];
      for (int q = 0; q < 3; q++) {
        s += Bb[r * n + 3 * S + q] * E[3 * q + cc];
        s += Bb[r * n + 3 * S + 3 + q] * E[9 + 3 * q + cc];
      }
    }
    a.fil_BA[36 * f + t] = s;
  }
}

// Schur complement of each swimmer and its inverse (swimmer.cpp:469-684)
// One workgroup of 256 per swimmer: the per-filament terms (the tether and
// clamp couplings and the filaments' B A blocks) are summed across the
// workgroup, then one lane assembles and inverts the 6x6.
__global__ __launch_bounds__(256) void k_schur_body(SimDev a, int nt) {
  __shared__ double sh[4 * 45];
  const int n = blockIdx.x;
  double *b = a.body + size_t(n) * BREC;
  const bool euler = nt < a.neuler;
  const double dtf = euler ? -a.dt : -2.0 * a.dt / 3.0;
  const d3 ub = ga3(b + BU);
  const m3 blc = mscale(-1.0, rcross(ub));
  m3 Qi;
  for (int q = 0; q < 9; q++) Qi.a[q] = b[BQI + q];
  // v[0..8]: sum over filaments of -rcross(teth) rcross(d) + 0.5 rcross(cl)
  //          - (rcross(u x cl) + blc rcross(cl)) / 12;  v[9..44]: sum of B A
  double v[45];
  for (int k = 0; k < 45; k++) v[k] = 0.0;
  const int f0 = a.sw_fil0[n], f1 = f0 + a.sw_nfil[n];
  for (int f = f0 + threadIdx.x; f < f1; f += 256) {
    const d3 d = mv(Qi, rowvec(a.fil_ref + 3 * f));
    const d3 teth = rowvec(a.teth + 9 * f), cl = rowvec(a.clamp + 9 * f);
    const d3 ucl = cross3(ub, cl);
    m3 t = madd(mscale(-1.0, mm(rcross(teth), rcross(d))),
                msub(mscale(0.5, rcross(cl)),
                     mscale(1.0 / 12.0, madd(rcross(ucl), mm(blc, rcross(cl))))));
    if (a.efm != 0.0) {  // swimmer.cpp:497-503
      const double dn = sqrt(dot3(d, d));
      const d3 v = ((0.5 * a.DL + dn) / dn) * d;
      t = madd(t, mm(rcross(a.efm * seg_t(a, f * a.S + a.S - 1)), rcross(v)));
    }
    for (int k = 0; k < 9; k++) v[k] += t.a[k];
    for (int k = 0; k < 36; k++) v[9 + k] += a.fil_BA[36 * f + k];
  }
  double r[45];
  wg_sums<45>(v, r, sh);
  if (threadIdx.x != 0) return;
  const m3 cbo = mscale(-1.0 / dtf, blc);
  double Sm[36];
  for (int k = 0; k < 36; k++) Sm[k] = (k % 7 == 0) ? 1.0 : 0.0;
  auto add33 = [&](const m3 &M) {
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) Sm[6 * (3 + i) + 3 + j] += M.a[3 * i + j];
  };
  add33(mscale(dtf, msub(mscale(1.0 / 12.0, mm(blc, cbo)), mscale(0.5, cbo))));
  m3 old;
  for (int k = 0; k < 9; k++) old.a[k] = r[k];
  double bm[36];
  rotate66(b + BN, Qi, bm);
  m3 bm33;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) bm33.a[3 * i + j] = bm[6 * (3 + i) + 3 + j];
  old = mm(bm33, old);
  add33(mscale(dtf, madd(msub(old, mscale(0.5, mm(blc, old))),
                         mscale(1.0 / 12.0, mm(mm(blc, blc), old)))));
  for (int k = 0; k < 36; k++) Sm[k] -= r[9 + k];
  if (!gj_small(Sm, 6)) atomicOr(&a.flags[F_SING], 1);
  for (int k = 0; k < 36; k++) b[BS + k] = Sm[k];
}

// -------------------------------------------------------------------------
// jacobian_inv_mult (swimmer.cpp:793-839)

__device__ __forceinline__ int fil_off(const SimDev &a, int f) {
  return 6 * a.S * f + 6 * a.fil_sw[f];
}
__device__ __forceinline__ int body_off(const SimDev &a, int n) {
  return 6 * a.S * (a.sw_fil0[n] + a.sw_nfil[n]) + 6 * n;
}

__global__ __launch_bounds__(64) void k_jim1(SimDev a, const double *in, const int *skip) {
  if (skip && *skip) return;
  __shared__ double sh[64];
  const int f = blockIdx.x;
  const int n = 6 * a.S;
  const double *x = in + fil_off(a, f);
  const double *Bb = a.Bblk + size_t(f) * 6 * n;
  for (int r = 0; r < 6; r++) {
    double s = 0;
    for (int c = threadIdx.x; c < n; c += 64) s += Bb[r * n + c] * x[c];
    s = block_sum<64>(s, sh);
    if (threadIdx.x == 0) a.fil_b2[6 * f + r] = s;
  }
}

// One workgroup of 256 per swimmer: the filaments' B_i in_i are summed
// across the workgroup, one lane applies S^-1.
__global__ __launch_bounds__(256) void k_jim_body(SimDev a, const double *in, double *out,
                                                  const int *skip) {
  if (skip && *skip) return;
  __shared__ double sh[4 * 6];
  const int n = blockIdx.x;
  const int o = body_off(a, n);
  double v[6] = {0, 0, 0, 0, 0, 0};
  const int f0 = a.sw_fil0[n], f1 = f0 + a.sw_nfil[n];
  for (int f = f0 + threadIdx.x; f < f1; f += 256)
    for (int k = 0; k < 6; k++) v[k] += a.fil_b2[6 * f + k];
  double sb[6];
  wg_sums<6>(v, sb, sh);
  if (threadIdx.x != 0) return;
  double b2[6];
  for (int k = 0; k < 6; k++) b2[k] = in[o + k] - sb[k];
  const double *Si = a.body + size_t(n) * BREC + BS;
  for (int r = 0; r < 6; r++) {
    double s = 0;
    for (int c = 0; c < 6; c++) s += Si[6 * r + c] * b2[c];
    out[o + r] = s;
  }
}

__global__ __launch_bounds__(256) void k_jim2(SimDev a, const double *in, double *out,
                                              const int *skip) {
  if (skip && *skip) return;
  __shared__ double tmp[6 * kMaxSeg];
  const int f = blockIdx.x;
  const int S = a.S, n = 6 * S;
  const int sw = a.fil_sw[f];
  const double *bo = out + body_off(a, sw);
  const d3 dx = {bo[0], bo[1], bo[2]}, du = {bo[3], bo[4], bo[5]};
  const double *b = a.body + size_t(sw) * BREC;
  m3 Qi;
  for (int q = 0; q < 9; q++) Qi.a[q] = b[BQI + q];
  const d3 d = mv(Qi, rowvec(a.fil_ref + 3 * f));
  const double dn = sqrt(dot3(d, d));
  const m3 Dm = rcross(((0.5 * a.DL + dn) / dn) * d);
  const d3 Ddu = mv(Dm, du);
  const double *E = a.ecb + 18 * f;
  m3 e1, e2;
  for (int q = 0; q < 9; q++) { e1.a[q] = E[q]; e2.a[q] = E[9 + q]; }
  const d3 a1 = mv(e1, du), a2 = mv(e2, du);
  const double *x = in + fil_off(a, f);
  for (int r = threadIdx.x; r < n; r += 256) {
    double Av = 0.0;
    if (r < 3 * S) {
      const int c = r % 3;
      Av = (c == 0 ? dx.x + Ddu.x : (c == 1 ? dx.y + Ddu.y : dx.z + Ddu.z));
    } else if (r < 3 * S + 3) {
      const int c = r - 3 * S;
      Av = c == 0 ? a1.x : (c == 1 ? a1.y : a1.z);
    } else if (r < 3 * S + 6) {
      const int c = r - 3 * S - 3;
      Av = c == 0 ? a2.x : (c == 1 ? a2.y : a2.z);
    }
    tmp[r] = x[r] - Av;
  }
  __syncthreads();
  const double *Ji = a.Jinv + size_t(f) * n * n;
  double *o = out + fil_off(a, f);
  for (int r = threadIdx.x; r < n; r += 256) {
    double s = 0;
    for (int c = 0; c < n; c++) s += Ji[size_t(r) * n + c] * tmp[c];
    o[r] = s;
  }
}

// -------------------------------------------------------------------------
// Broyden (broyden_solver.cpp:26-130)

// upd = -gamma*upd - sum_i d[i] C_i   (find_update, and the new C column)
__global__ void k_broyden_combine(double *out, const double *in, double gamma, const double *C,
                                  size_t ldc, const double *d, int k, int n, const int *skip) {
  if (skip && *skip) return;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  double s = -gamma * in[e];
  for (int i = 0; i < k; i++) s -= d[i] * C[ldc * i + e];
  out[e] = s;
}

__global__ void k_scale2(double *x, double *y, int n, const double *scal, int o,
                         const int *skip) {
  if (skip && *skip) return;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) { x[e] /= scal[o]; y[e] /= scal[o]; }
}

__global__ void k_set_step(SimDev a, int nt, double t) {
  a.step_nt[0] = nt;
  a.step_t[0] = t;
}

// Gates of a captured Broyden iteration.  After the first GMRES chunk:
// hold the rest of the iteration unless the solve has converged.  After the
// residual: hold the Broyden update unless the residual is finite and too
// large (the host then takes the iteration as converged or diverged).
__global__ void k_hold_unsolved(int *flags) { flags[F_HOLD] = flags[F_DONE] ? 0 : 1; }
__global__ void k_hold_unless_big(int *flags) {
  flags[F_HOLD] = (flags[F_HOLD] || !flags[F_BIG] || flags[F_NAN]) ? 1 : 0;
}

// -------------------------------------------------------------------------
// state update (swimmer.cpp:847-871, rigid_body.cpp:354-381, filament.cpp:1541-1575)

__global__ void k_body_update(SimDev a, const double *upd) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= a.Nsw) return;
  double *b = a.body + size_t(n) * BREC;
  const double *u = upd + body_off(a, n);
  for (int k = 0; k < 3; k++) { b[BX + k] += u[k]; b[BU + k] += u[3 + k]; }
  sa4(b + BQ, qmul(lie_exp(ga3(b + BU)), ga4(b + BQM1)));
}

// One wave per filament: lane 0 takes the base from the body (accept_base)
// and the tether and clamp multipliers, lane j > 0 its multiplier and segment
// update; the tangents then go through LDS to lane 0, which lays the robot
// arm out in segment order.  (The reference's accept_state_from_rigid_body
// also lays the arm out before the updates; the layout after them rewrites
// every position, so it is done once.)
__global__ __launch_bounds__(kFilLanes) void k_fil_update(SimDev a, const double *upd) {
  __shared__ double stn[3 * kFilLanes];
  const int f = blockIdx.x, j = threadIdx.x;
  const int S = a.S, g0 = f * S;
  const double *u = upd + fil_off(a, f);
  if (j == 0) {
    accept_base(a, f, g0);
    double *T = a.teth + 9 * f;
    for (int k = 0; k < 3; k++) T[k] += u[k];
    double *Cl = a.clamp + 9 * f;
    for (int k = 0; k < 3; k++) Cl[k] += u[3 * S + k];
  } else if (j < S) {
    double *L = a.lam + size_t(f) * 9 * S;
    for (int k = 0; k < 3; k++) L[3 * (j - 1) + k] += u[3 * j + k];
    seg_update(a, g0 + j, rowvec(u + 3 * (S + j)));
  }
  if (j < S) sa3(stn + 3 * j, seg_t(a, g0 + j));
  __syncthreads();
  if (j == 0) robot_arm(a, g0, stn);
}

}  // namespace

// ==========================================================================
// host side

static SimHost *sim_of(gmi_ctx *c) {
  return c->extra_kind == kExtraSim ? static_cast<SimHost *>(c->extra) : nullptr;
}

static void sim_release(gmi_ctx *c) {
  SimHost *s = sim_of(c);
  if (!s) return;
  s->drop_graphs();
  learn_release(s);
  dense_release(s);
  for (void *p : s->owned) (void)hipFree(p);
  if (s->h_flags) (void)hipHostFree(s->h_flags);
  if (s->h_scal) (void)hipHostFree(s->h_scal);
  delete s;
}

template <typename T>
static int dalloc(gmi_ctx *c, SimHost *s, T **p, size_t n) {
  void *q = nullptr;
  int rc = hip_check(c, hipMalloc(&q, sizeof(T) * std::max<size_t>(n, 1)), "hipMalloc sim");
  if (rc) return rc;
  s->owned.push_back(q);
  *p = static_cast<T *>(q);
  return hip_check(c, hipMemsetAsync(q, 0, sizeof(T) * std::max<size_t>(n, 1), c->stream),
                   "memset sim");
}

template <typename T>
static int upload(gmi_ctx *c, T *dst, const T *src, size_t n) {
  if (n == 0) return GMI_OK;
  int rc = hip_check(c, hipMemcpyAsync(dst, src, sizeof(T) * n, hipMemcpyHostToDevice, c->stream),
                     "upload");
  if (rc) return rc;
  return hip_check(c, hipStreamSynchronize(c->stream), "sync");
}

static int read_flags(gmi_ctx *c, SimHost *s) { return krylov_read_flags(c, s); }

// system_matrix_mult: out = A z (A as in mobility_solver.cpp:905-964).
// mode 0: FP64 pairs, 1: FP32 pairs, 2: both launched, each gated by its
// device flag (F_SKIP64 / F_SKIP32) so the precision switch needs no sync.
// The tile shares of -K^T z are in s->zpart when zpart_ready (z came from
// the preconditioner); otherwise they are computed first.
// The blob-blob pair sums of out = A z (split partials left for the K products).
static int sys_pairs(gmi_ctx *c, SimHost *s, double *z, int mode, const int *done) {
  SimDev &d = s->d;
  PairArgs pa;
  const WorkList &wl = c->lists[kBB];
  pa.Xs = c->Xs; pa.Fs = c->Fs; pa.lds = c->lds;
  pa.Xb = c->Xb; pa.Fb = z; pa.ldb = c->ldb;
  pa.Vs = c->Vs; pa.Vb = c->Vb;
  pa.part = c->part; pa.tiles = wl.d_tiles; pa.work = wl.d_work;
  pa.mask = kBB; pa.accumulate = 0;
  pa.a_s = c->as; pa.a_b = c->ab; pa.mu = c->mu;
  int rc;
  // split partials are summed by k_kprod (its tiles are the list's tiles)
  if (mode == 2) {
    // one launch: FP32 once the relaxed switch cleared F_SKIP32, else FP64
    pa.skip = d.flags + F_DONE;
    pa.sel32 = d.flags + F_SKIP32;
    if ((rc = hip_check(c, launch_pair_dyn(pa, wl.nwork, 0, c->stream), "pair bb"))) return rc;
  } else {
    pa.skip = done;
    if ((rc = hip_check(c, launch_pair(pa, wl.nwork, 0, mode == 1, c->stream), "pair bb")))
      return rc;
  }
  return GMI_OK;
}

static int sys_mult(gmi_ctx *c, SimHost *s, double *z, double *out, int mode, const int *done,
                    bool zpart_ready) {
  SimDev &d = s->d;
  if (!zpart_ready)
    hipLaunchKernelGGL(k_precond_kt, dim3(s->nbtiles), dim3(256), 0, c->stream, d, s->btiles,
                       s->sw_tile0, z, s->scal, -1, s->zpart, done);
  int rc;
  if ((rc = sys_pairs(c, s, z, mode, done))) return rc;
  const WorkList &wl = c->lists[kBB];
  hipLaunchKernelGGL(k_kprod, dim3(s->nbtiles), dim3(256), 0, c->stream, d, s->btiles,
                     s->sw_tile0, s->sw_ntile, wl.d_tiles, c->part, z, c->Vb, out, s->zpart,
                     done);
  return launch_ok(c, "kprod");
}

// apply_preconditioner: out = P^-1 in.  norm_slot >= 0: `in` is first
// divided by scal[norm_slot] in place (the previous Arnoldi vector's norm).
static int precond(gmi_ctx *c, SimHost *s, double *in, double *out, const int *skip,
                   int norm_slot) {
  const SimDev &d = s->d;
  hipLaunchKernelGGL(k_precond_kt, dim3(s->nbtiles), dim3(256), 0, c->stream, d, s->btiles,
                     s->sw_tile0, in, s->scal, norm_slot, s->tpart, skip);
  hipLaunchKernelGGL(k_precond_blobs, dim3(s->nbtiles), dim3(256), 0, c->stream, d, s->btiles,
                     s->sw_tile0, s->sw_ntile, s->tpart, s->rot, in, out, s->zpart, skip);
  return launch_ok(c, "precond");
}

// The rigid-body system as the Krylov engine's operator.
struct SimOp {
  gmi_ctx *c;
  SimHost *s;
  int mult(double *z, double *out, int mode, const int *done, bool zpart_ready) {
    // the body-frame preconditioner leaves no K^T shares behind
    return sys_mult(c, s, z, out, mode, done, zpart_ready && !s->dense);
  }
  // the exact body-frame preconditioner: FP64 Krylov products, and a start
  // from zero (the inverse is exact, so a warm start saves no step)
  bool fp64_krylov() const { return s->dense != nullptr; }
  bool zero_start() const { return s->dense != nullptr; }
  int precond(double *in, double *out, const int *skip, int norm_slot) {
    if (s->dense) return dense_apply(c, s, in, out, skip, norm_slot);
    return gmi::precond(c, s, in, out, skip, norm_slot);
  }
  // the reference's preconditioner: whole Arnoldi steps on 64-blob chunks
  bool fused() const { return s->dense == nullptr; }
  int fused_parts() const { return s->nchunks; }
  int fused_step(KrylovState *, int it, bool chunk_start, bool mixed, const GivensArgs &g) {
    SimDev &d = s->d;
    const size_t ldq = size_t(d.npad);
    double *qprev = s->Qk + ldq * (it - 1), *w = s->Qk + ldq * it;
    const int nc = s->nchunks;
    if (it == 1)  // -K^T q_0 shares (q_0 normalised by the GMRES start)
      hipLaunchKernelGGL(k_precond_kt, dim3(nc), dim3(256), 0, c->stream, d, s->chunks,
                         s->sw_chunk0, qprev, s->scal, -1, s->ch_kpart, s->flags);
    hipLaunchKernelGGL(k_arnoldi_precond, dim3(nc), dim3(256), 0, c->stream, d, s->chunks,
                       s->sw_chunk0, s->sw_nchunk, s->ch_kpart, s->part, nc, s->rot, qprev, s->z,
                       s->ch_zpart, g, it - 1, it == 1 ? 2 : (chunk_start ? 0 : 1));
    int rc;
    if ((rc = launch_ok(c, "arnoldi precond"))) return rc;
    if ((rc = sys_pairs(c, s, s->z, mixed ? 2 : 0, s->flags))) return rc;
    hipEvent_t gs_stop = nullptr;
    if ((rc = prof_gs_start(c, it, d.npad, s->gs_passes, &gs_stop))) return rc;
    double *pp1 = s->pp, *pp2 = s->pp + s->pp_half;
    const WorkList &wl = c->lists[kBB];
    hipLaunchKernelGGL(k_kprod_dots, dim3(nc), dim3(kGsThreads), 0, c->stream, d, s->chunks, s->ch_tile,
                       s->sw_chunk0, s->sw_nchunk, wl.d_tiles, c->part, s->z, c->Vb, w,
                       s->ch_zpart, s->flags, s->Qk, ldq, it, it > 1 ? 1 : 0, s->scal, pp1, nc);
    // beyond kWideChunks chunks every workgroup re-summing every column's
    // partials would cost more than one launch that sums each column once
    const bool wide = nc > kWideChunks;
    if (wide)
      hipLaunchKernelGGL(k_colsum_wide, dim3(it), dim3(256), 0, c->stream, pp1, nc, s->h1,
                         s->flags);
    if (s->gs_passes == 1) {
      // one pass, as the reference's modified Gram-Schmidt takes one
      hipLaunchKernelGGL(k_gs_chunk<2>, dim3(nc), dim3(kGsThreads), 0, c->stream, d, s->chunks,
                         s->sw_chunk0, s->Qk, ldq, it, pp1, s->h1, w, nullptr, nc, s->part,
                         s->ch_kpart, s->flags, int(wide), s->h2);
    } else {
      hipLaunchKernelGGL(k_gs_chunk<1>, dim3(nc), dim3(kGsThreads), 0, c->stream, d, s->chunks,
                         s->sw_chunk0, s->Qk, ldq, it, pp1, s->h1, w, pp2, nc, nullptr, nullptr,
                         s->flags, int(wide), nullptr);
      if (wide)
        hipLaunchKernelGGL(k_colsum_wide, dim3(it), dim3(256), 0, c->stream, pp2, nc, s->h2,
                           s->flags);
      hipLaunchKernelGGL(k_gs_chunk<2>, dim3(nc), dim3(kGsThreads), 0, c->stream, d, s->chunks,
                         s->sw_chunk0, s->Qk, ldq, it, pp2, s->h2, w, nullptr, nc, s->part,
                         s->ch_kpart, s->flags, int(wide), nullptr);
    }
    if (gs_stop && (rc = hip_check(c, hipEventRecord(gs_stop, c->stream), "event"))) return rc;
    return launch_ok(c, "arnoldi step");
  }
};

// The fitted GMRES initial guess of the coming solve (history on): the
// right-hand side's fit onto the history, the combination of the history's
// solutions as x (an empty history leaves the previous solution).
static int hist_guess(gmi_ctx *c, SimHost *s) {
  if (s->hist_m == 0) return GMI_OK;
  const SimDev &d = s->d;
  const size_t ldh = size_t(d.npad);
  hipLaunchKernelGGL(k_hist_dots, dim3(s->nbtiles), dim3(256), 0, c->stream, d, s->btiles,
                     s->sw_tile0, s->rhs, s->hist_bnew, s->hist_b, ldh, s->hist_st, s->hist_part);
  hipLaunchKernelGGL(k_hist_solve, dim3(nblk(d.Nsw, 64)), dim3(64), 0, c->stream, d, s->sw_tile0,
                     s->sw_ntile, s->hist_part, s->hist_G, s->hist_Gw, s->hist_st, s->hist_coef,
                     s->hist_gsave);
  hipLaunchKernelGGL(k_hist_combine, dim3(s->nbtiles), dim3(256), 0, c->stream, d, s->btiles,
                     s->sw_tile0, s->hist_x, ldh, s->hist_st, s->hist_coef, s->x);
  return launch_ok(c, "fitted guess");
}

// The solved system into the history (history on); skip gates the kernels.
static int hist_push(gmi_ctx *c, SimHost *s, const int *skip) {
  if (s->hist_m == 0) return GMI_OK;
  const SimDev &d = s->d;
  hipLaunchKernelGGL(k_hist_push, dim3(s->nbtiles), dim3(256), 0, c->stream, d, s->btiles,
                     s->sw_tile0, s->hist_bnew, s->x, s->hist_b, s->hist_x, size_t(d.npad),
                     s->hist_st, s->hist_G, s->hist_gsave, skip);
  hipLaunchKernelGGL(k_hist_advance, dim3(1), dim3(1), 0, c->stream, s->hist_st, s->hist_m, skip);
  return launch_ok(c, "history push");
}

// Rigid-body solve (mobility_solver.cpp:970-1354): the body matrices are
// rotated once, then GMRES on the rigid-body operator, its initial guess
// fitted on the history when that is on.
static int solve_linear_system(gmi_ctx *c, SimHost *s, int *its) {
  SimDev &d = s->d;
  hipLaunchKernelGGL(k_rotate_body, dim3(nblk(d.Nsw, 64)), dim3(64), 0, c->stream, d, s->rot);
  int rc;
  if ((rc = launch_ok(c, "rotate body"))) return rc;
  if ((rc = hist_guess(c, s))) return rc;
  SimOp op{c, s};
  if ((rc = krylov_solve(c, s, op, d.npad, std::min(d.maxit, 3 * d.Nb + 6 * d.Nsw), s->precision,
                         s->p.gmres_tol, its)))
    return rc;
  return hist_push(c, s, nullptr);
}

// read_positions_and_forces (mobility_solver.cpp:132-192)
static int read_positions_and_forces(gmi_ctx *c, SimHost *s) {
  SimDev &d = s->d;
  hipLaunchKernelGGL(k_blob_positions, dim3(nblk(d.Nb, 256)), dim3(256), 0, c->stream, d);
  int rc;
  if ((rc = launch_ok(c, "blob positions"))) return rc;
  if ((rc = hip_check(c, hipMemcpyAsync(c->Fs, d.Fint, sizeof(double) * 6 * c->lds,
                                        hipMemcpyDeviceToDevice, c->stream), "forces to device")))
    return rc;
  if ((rc = gmi_apply_interparticle_forces(c, s->p.barrier_K))) return rc;
  hipLaunchKernelGGL(k_body_repulsion, dim3(d.Nsw), dim3(256), 0, c->stream, d);
  return launch_ok(c, "body repulsion");
}

// The blob–blob solve's right-hand side from the current state: blob
// positions, barrier, segment-segment and blob<-segment sums.
static int velocities_before_solve(gmi_ctx *c, SimHost *s) {
  int rc;
  if ((rc = read_positions_and_forces(c, s))) return rc;
  const bool fp32 = s->precision == GMI_FP32_KAHAN;
  if ((rc = run_pairs(c, kSS | kBS, 0, fp32, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr)))
    return rc;
  hipLaunchKernelGGL(k_assemble_rhs, dim3(nblk(s->d.npad, 256)), dim3(256), 0, c->stream, s->d,
                     s->rhs);
  return launch_ok(c, "rhs");
}

// segment (v, w) += the blob forces' contribution (after the solve)
static int velocities_after_solve(gmi_ctx *c, SimHost *s, const int *skip) {
  return run_pairs(c, kSB, 1, s->precision == GMI_FP32_KAHAN, nullptr, s->x, nullptr, nullptr,
                   nullptr, nullptr, skip);
}

static int compute_velocities(gmi_ctx *c, SimHost *s, int *its) {
  int rc;
  if ((rc = velocities_before_solve(c, s))) return rc;
  int it = 0;
  if ((rc = solve_linear_system(c, s, &it))) return rc;
  s->last_gmres = it;
  if (its) *its = it;
  return velocities_after_solve(c, s, nullptr);
}

// the step's time and index for the kernels that read them on the device
static int set_step(gmi_ctx *c, SimHost *s, int nt) {
  hipLaunchKernelGGL(k_set_step, dim3(1), dim3(1), 0, c->stream, s->d, nt,
                     (nt + 1) * s->p.dt);
  return launch_ok(c, "step");
}

// Residual into dst and its flags (F_BIG, F_NAN) on the device.
static int enqueue_errors(gmi_ctx *c, SimHost *s, double *dst, const int *skip) {
  SimDev &d = s->d;
  hipLaunchKernelGGL(k_set_ints, dim3(1), dim3(64), 0, c->stream, d.flags + F_BIG, 2, 0, skip);
  hipLaunchKernelGGL(k_errors, dim3(nblk(d.Ns + d.Nsw, 256)), dim3(256), 0, c->stream, d, dst,
                     s->x + 3 * d.ldb, skip);
  return launch_ok(c, "errors");
}

static int compute_errors(gmi_ctx *c, SimHost *s, double *dst, int *big, int *bad) {
  int rc;
  if ((rc = enqueue_errors(c, s, dst, nullptr))) return rc;
  if ((rc = read_flags(c, s))) return rc;
  if (big) *big = s->h_flags[F_BIG];
  if (bad) *bad = s->h_flags[F_NAN];
  return GMI_OK;
}

static int jacobian_inv_mult(gmi_ctx *c, SimHost *s, const double *in, double *out,
                             const int *skip) {
  SimDev &d = s->d;
  hipLaunchKernelGGL(k_jim1, dim3(d.Nf), dim3(64), 0, c->stream, d, in, skip);
  hipLaunchKernelGGL(k_jim_body, dim3(d.Nsw), dim3(256), 0, c->stream, d, in, out, skip);
  hipLaunchKernelGGL(k_jim2, dim3(d.Nf), dim3(256), 0, c->stream, d, in, out, skip);
  return launch_ok(c, "jacobian_inv_mult");
}

}  // namespace gmi

using namespace gmi;

extern "C" {

int gmi_sim_create(gmi_ctx *c, const gmi_sim_params *p) {
  if (!c || !p || !p->nfil || !p->blob_refs || !p->fil_refs) return GMI_ERR_ARG;
  if (c->extra) return fail(c, GMI_ERR_STATE, "context already has a sim");
  if (c->wall) return fail(c, GMI_ERR_STATE, "no time stepping above a wall");
  const int S = c->nseg_fil;
  if (S < 2 || S > kMaxSeg) return fail(c, GMI_ERR_ARG, "NSEG must be in [2, 64]");
  int Nf = 0;
  for (int n = 0; n < c->nsw; n++) {
    if (p->nfil[n] < 1 || p->nfil[n] * S != c->nseg[n])
      return fail(c, GMI_ERR_ARG, "nfil*NSEG must match the context's segment counts");
    if (c->nblob[n] < 1) return fail(c, GMI_ERR_ARG, "every swimmer needs a blob body");
    Nf += p->nfil[n];
  }
  if (!(p->dt > 0) || !(p->tol > 0) || !(p->gmres_tol > 0) || p->max_broyden_iter < 1 ||
      p->gmres_max_iter < 1 || p->gmres_max_iter > kMaxCols - 1 || !(p->sphere_radius > 0) || !(p->curvature_period > 0))
    return fail(c, GMI_ERR_ARG, "bad sim parameters");
  if (p->sphere_radii)
    for (int n = 0; n < c->nsw; n++)
      if (!(p->sphere_radii[n] > 0)) return fail(c, GMI_ERR_ARG, "sphere_radii must be positive");
  if (!std::isfinite(p->end_force)) return fail(c, GMI_ERR_ARG, "end_force must be finite");
  if (p->perturbation && S < 13)
    return fail(c, GMI_ERR_ARG, "the perturbations act on segments 4 and 12: NSEG >= 13");
  SimHost *s = new SimHost();
  s->p = *p;
  // the caller's arrays need not outlive this call: keep copies of what is
  // read later (blob and filament references), drop the other pointers
  {
    size_t nbr = 0, nfr = 0;
    for (int n = 0; n < c->nsw; n++) { nbr += 3 * size_t(c->nblob[n]); nfr += 3 * size_t(p->nfil[n]); }
    s->blob_refs_h.assign(p->blob_refs, p->blob_refs + nbr);
    s->fil_refs_h.assign(p->fil_refs, p->fil_refs + nfr);
    s->p.blob_refs = s->blob_refs_h.data();
    s->p.fil_refs = s->fil_refs_h.data();
    s->p.nfil = nullptr;
    s->p.body_x = nullptr;
    s->p.fil_phase = nullptr;
    s->p.sphere_radii = nullptr;
    s->p.perturbation = nullptr;
  }
  s->precision = c->precision;
  c->extra = s;
  c->extra_release = sim_release;
  c->extra_kind = kExtraSim;
  SimDev &d = s->d;
  d.S = S; d.Nf = Nf; d.Nsw = c->nsw; d.Ns = c->Ns; d.Nb = c->Nb;
  d.lds = c->lds; d.ldb = c->ldb;
  d.nbroy = 6 * (Nf * S + c->nsw);
  d.npad = 3 * c->ldb + 6 * c->nsw;
  d.maxit = p->gmres_max_iter;
  d.DL = 2.2 * c->as; d.mu = c->mu; d.as = c->as; d.ab = c->ab;
  d.kb = p->kb; d.kt = p->kt; d.dt = p->dt; d.tol = p->tol; d.gamma = p->jacobian_confidence;
  d.gtol = p->gmres_tol; d.neuler = p->num_euler_steps;
  d.camp = p->curvature_amp; d.cper = p->curvature_period;
  d.efm = p->end_force;
  d.pert = nullptr;
  d.X = c->Xs; d.Fs = c->Fs; d.Vs = c->Vs; d.Xb = c->Xb; d.Vb = c->Vb; d.Frep = c->Frep;
  int rc;
  const size_t ls = size_t(d.lds), nn = size_t(6 * S) * (6 * S);
  int *ip;
  if ((rc = dalloc(c, s, &ip, d.Nsw))) return rc; d.sw_fil0 = ip;
  if ((rc = dalloc(c, s, &ip, d.Nsw))) return rc; d.sw_nfil = ip;
  if ((rc = dalloc(c, s, &ip, d.Nsw))) return rc; d.sw_blob0 = ip;
  if ((rc = dalloc(c, s, &ip, d.Nsw))) return rc; d.sw_nblob = ip;
  if ((rc = dalloc(c, s, &ip, Nf))) return rc; d.fil_sw = ip;
  if ((rc = dalloc(c, s, &ip, d.Nb))) return rc; d.blob_sw = ip;
  double *swR;
  if ((rc = dalloc(c, s, &swR, d.Nsw))) return rc; d.sw_R = swR;
  if ((rc = dalloc(c, s, &d.flags, F_N))) return rc;
  s->flags = d.flags;
  if ((rc = dalloc(c, s, &d.step_t, 1)) || (rc = dalloc(c, s, &d.step_nt, 1))) return rc;
  s->cols = d.maxit;
  if ((rc = dalloc(c, s, &d.Xm1, 3 * ls))) return rc;
  if ((rc = dalloc(c, s, &d.Xm2, 3 * ls))) return rc;
  if ((rc = dalloc(c, s, &d.Q, 4 * ls))) return rc;
  if ((rc = dalloc(c, s, &d.Qm1, 4 * ls))) return rc;
  if ((rc = dalloc(c, s, &d.U, 3 * ls))) return rc;
  if ((rc = dalloc(c, s, &d.Um1, 3 * ls))) return rc;
  if ((rc = dalloc(c, s, &d.Fint, 6 * ls))) return rc;
  if ((rc = dalloc(c, s, &d.fil_ref, 3 * size_t(Nf)))) return rc;
  if ((rc = dalloc(c, s, &d.phase, Nf))) return rc;
  if ((rc = dalloc(c, s, &d.lam, size_t(Nf) * 9 * S))) return rc;
  if ((rc = dalloc(c, s, &d.teth, 9 * size_t(Nf)))) return rc;
  if ((rc = dalloc(c, s, &d.clamp, 9 * size_t(Nf)))) return rc;
  if ((rc = dalloc(c, s, &d.fil_bf, 6 * size_t(Nf)))) return rc;
  if ((rc = dalloc(c, s, &d.body, size_t(d.Nsw) * BREC))) return rc;
  if ((rc = dalloc(c, s, &d.blob_ref, 3 * size_t(d.ldb)))) return rc;
  if ((rc = dalloc(c, s, &d.Jw, size_t(Nf) * nn))) return rc;
  if ((rc = dalloc(c, s, &d.Jinv, size_t(Nf) * nn))) return rc;
  if ((rc = dalloc(c, s, &d.gj_redo, size_t(Nf)))) return rc;
  if ((rc = dalloc(c, s, &d.ecb, 18 * size_t(Nf)))) return rc;
  if ((rc = dalloc(c, s, &d.Bblk, size_t(Nf) * 36 * S))) return rc;
  if ((rc = dalloc(c, s, &d.fil_BA, 36 * size_t(Nf)))) return rc;
  if ((rc = dalloc(c, s, &d.fil_b2, 6 * size_t(Nf)))) return rc;
  const size_t np = size_t(d.npad);
  const int maxit = d.maxit;
  if ((rc = dalloc(c, s, &s->x, np))) return rc;
  if ((rc = dalloc(c, s, &s->rhs, np))) return rc;
  if ((rc = dalloc(c, s, &s->Qk, np * (maxit + 1)))) return rc;
  if ((rc = dalloc(c, s, &s->H, size_t(maxit + 1) * (maxit + 1)))) return rc;
  if ((rc = dalloc(c, s, &s->cs, maxit + 1))) return rc;
  if ((rc = dalloc(c, s, &s->sn, maxit + 1))) return rc;
  if ((rc = dalloc(c, s, &s->beta, maxit + 2))) return rc;
  if ((rc = dalloc(c, s, &s->bpre, maxit + 2))) return rc;
  if ((rc = dalloc(c, s, &s->h1, maxit + 1))) return rc;
  if ((rc = dalloc(c, s, &s->h2, maxit + 1))) return rc;
  if ((rc = dalloc(c, s, &s->y, maxit + 1))) return rc;
  if ((rc = dalloc(c, s, &s->z, np))) return rc;
  if ((rc = dalloc(c, s, &s->tmp, np))) return rc;
  if ((rc = dalloc(c, s, &s->scal, 16))) return rc;
  // blob chunks of the fused Arnoldi steps (built below): up to 64 blobs of
  // one swimmer each
  size_t ntiles_blob = 0;
  for (int n = 0; n < d.Nsw; n++) ntiles_blob += (c->nblob[n] + kChunk - 1) / kChunk;
  if ((rc = dalloc(c, s, &s->part, std::max<size_t>(nblk(d.npad, 256), ntiles_blob)))) return rc;
  // column dot partials: two halves (Gram-Schmidt passes 1 and 2), per
  // row chunk (generic kernels, Broyden D) or per blob tile (fused steps)
  s->pp_half = std::max(std::max(size_t(maxit + 1) * dot_shape(d.npad).nb,
                                 size_t(maxit + 1) * ntiles_blob),
                        size_t(p->max_broyden_iter) * dot_shape(d.nbroy).nb);
  if ((rc = dalloc(c, s, &s->pp, 2 * s->pp_half))) return rc;
  if ((rc = dalloc(c, s, &s->rot, 72 * size_t(d.Nsw)))) return rc;
  const size_t nbr = size_t(d.nbroy);
  if ((rc = dalloc(c, s, &s->err, nbr))) return rc;
  if ((rc = dalloc(c, s, &s->nerr, nbr))) return rc;
  s->err_first = s->err;
  if ((rc = dalloc(c, s, &s->upd, nbr))) return rc;
  if ((rc = dalloc(c, s, &s->C, nbr * p->max_broyden_iter))) return rc;
  // D holds MAX_BROYDEN_ITER columns plus two GMRES-refinement scratch vectors
  if ((rc = dalloc(c, s, &s->D, nbr * p->max_broyden_iter + 2 * std::max(nbr, np)))) return rc;
  if ((rc = dalloc(c, s, &s->bd, p->max_broyden_iter + 1))) return rc;
  if ((rc = hip_check(c, hipHostMalloc((void **)&s->h_flags, sizeof(int) * F_N,
                                       hipHostMallocDefault), "pinned flags")))
    return rc;
  if ((rc = hip_check(c, hipHostMalloc((void **)&s->h_scal, sizeof(double) * 16,
                                       hipHostMallocDefault), "pinned scal")))
    return rc;
  // ---- host-side initial state (swimmer.cpp:160-237, filament.cpp:197-340)
  std::vector<int> fil0(d.Nsw), nfv(d.Nsw), blob0(d.Nsw), nblv(d.Nsw), fil_sw(Nf), blob_sw(d.Nb);
  int fo = 0;
  for (int n = 0; n < d.Nsw; n++) {
    fil0[n] = fo; nfv[n] = p->nfil[n];
    blob0[n] = c->blob_off[n]; nblv[n] = c->nblob[n];
    for (int i = 0; i < p->nfil[n]; i++) fil_sw[fo + i] = n;
    for (int m = 0; m < c->nblob[n]; m++) blob_sw[c->blob_off[n] + m] = n;
    fo += p->nfil[n];
  }
  std::vector<double> body(size_t(d.Nsw) * BREC, 0.0);
  for (int n = 0; n < d.Nsw; n++) {
    double *b = &body[size_t(n) * BREC];
    for (int k = 0; k < 3; k++) b[BX + k] = b[BXM1 + k] = b[BXM2 + k] = p->body_x ? p->body_x[3 * n + k] : 0.0;
    b[BQ] = b[BQM1] = 1.0;
    for (int k = 0; k < 6; k++) { b[BN + 7 * k] = 1.0; b[BK + 7 * k] = 1.0; }
    b[BQI] = b[BQI + 4] = b[BQI + 8] = 1.0;
  }
  std::vector<double> bref(3 * size_t(d.ldb), 0.0);
  for (int m = 0; m < d.Nb; m++)
    for (int k = 0; k < 3; k++) bref[k * d.ldb + m] = p->blob_refs[3 * m + k];
  std::vector<double> X(3 * ls, 0.0), Qv(4 * ls, 0.0), phase(Nf, 0.0);
  for (int f = 0; f < Nf; f++) {
    phase[f] = p->fil_phase ? p->fil_phase[f] : 0.0;
    const double *r = p->fil_refs + 3 * f;
    const double rn = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    const double dd[3] = {r[0] / rn, r[1] / rn, r[2] / rn};
    // rotation taking the x axis to dd: sqrt of (dd.x, x cross dd)
    double qs = dd[0], qx = 0.0, qy = -dd[2], qz = dd[1];
    if (qs < -0.999999999) { qs = 0; qx = 0; qy = 0; qz = 1; }
    else {
      qs = std::sqrt(0.5 * (1.0 + qs));
      const double t = 2.0 * qs;
      qx /= t; qy /= t; qz /= t;
    }
    const double *bx = &body[size_t(fil_sw[f]) * BREC + BX];
    for (int j = 0; j < S; j++) {
      const int g = f * S + j;
      for (int k = 0; k < 3; k++) X[k * ls + g] = bx[k] + r[k] + j * d.DL * dd[k];
      Qv[g] = qs; Qv[ls + g] = qx; Qv[2 * ls + g] = qy; Qv[3 * ls + g] = qz;
    }
  }
  if ((rc = upload(c, const_cast<int *>(d.sw_fil0), fil0.data(), d.Nsw))) return rc;
  if ((rc = upload(c, const_cast<int *>(d.sw_nfil), nfv.data(), d.Nsw))) return rc;
  if ((rc = upload(c, const_cast<int *>(d.sw_blob0), blob0.data(), d.Nsw))) return rc;
  if ((rc = upload(c, const_cast<int *>(d.sw_nblob), nblv.data(), d.Nsw))) return rc;
  if ((rc = upload(c, const_cast<int *>(d.fil_sw), fil_sw.data(), Nf))) return rc;
  if ((rc = upload(c, const_cast<int *>(d.blob_sw), blob_sw.data(), d.Nb))) return rc;
  {
    std::vector<double> R(d.Nsw, p->sphere_radius);
    if (p->sphere_radii) R.assign(p->sphere_radii, p->sphere_radii + d.Nsw);
    if ((rc = upload(c, const_cast<double *>(d.sw_R), R.data(), d.Nsw))) return rc;
  }
  if ((rc = upload(c, d.body, body.data(), body.size()))) return rc;
  if ((rc = upload(c, d.blob_ref, bref.data(), bref.size()))) return rc;
  if ((rc = upload(c, d.fil_ref, p->fil_refs, 3 * size_t(Nf)))) return rc;
  if ((rc = upload(c, d.phase, phase.data(), phase.size()))) return rc;
  if (p->perturbation) {
    double *pp;
    if ((rc = dalloc(c, s, &pp, 6 * size_t(Nf)))) return rc;
    if ((rc = upload(c, pp, p->perturbation, 6 * size_t(Nf)))) return rc;
    d.pert = pp;
  }
  if ((rc = upload(c, d.X, X.data(), X.size()))) return rc;
  if ((rc = upload(c, d.Xm1, X.data(), X.size()))) return rc;
  if ((rc = upload(c, d.Xm2, X.data(), X.size()))) return rc;
  if ((rc = upload(c, d.Q, Qv.data(), Qv.size()))) return rc;
  if ((rc = upload(c, d.Qm1, Qv.data(), Qv.size()))) return rc;
  s->nfil.assign(p->nfil, p->nfil + d.Nsw);
  // blob tiles of <= 256 blobs inside one swimmer
  std::vector<BTile> bt;
  std::vector<int> t0(d.Nsw), nt(d.Nsw);
  for (int n = 0; n < d.Nsw; n++) {
    t0[n] = int(bt.size());
    for (int m = 0; m < c->nblob[n]; m += 256)
      bt.push_back({n, c->blob_off[n] + m, std::min(256, c->nblob[n] - m)});
    nt[n] = int(bt.size()) - t0[n];
  }
  s->nbtiles = int(bt.size());
  BTile *dbt;
  if ((rc = dalloc(c, s, &dbt, bt.size()))) return rc;
  s->btiles = dbt;
  if ((rc = upload(c, dbt, bt.data(), bt.size()))) return rc;
  int *i0, *i1;
  if ((rc = dalloc(c, s, &i0, d.Nsw))) return rc;
  if ((rc = dalloc(c, s, &i1, d.Nsw))) return rc;
  if ((rc = upload(c, i0, t0.data(), d.Nsw))) return rc;
  if ((rc = upload(c, i1, nt.data(), d.Nsw))) return rc;
  s->sw_tile0 = i0;
  s->sw_ntile = i1;
  {
    // 64-blob chunks of the tiles, in tile order
    std::vector<BTile> ch;
    std::vector<int> ct, c0(d.Nsw), cn(d.Nsw);
    for (int n = 0; n < d.Nsw; n++) {
      c0[n] = int(ch.size());
      for (int q = t0[n]; q < t0[n] + nt[n]; q++)
        for (int o = 0; o < bt[q].n; o += kChunk) {
          ch.push_back({n, bt[q].m0 + o, std::min(kChunk, bt[q].n - o)});
          ct.push_back(q);
        }
      cn[n] = int(ch.size()) - c0[n];
    }
    s->nchunks = int(ch.size());
    BTile *dch;
    int *dct, *dc0, *dcn;
    if ((rc = dalloc(c, s, &dch, ch.size()))) return rc;
    if ((rc = dalloc(c, s, &dct, ct.size()))) return rc;
    if ((rc = dalloc(c, s, &dc0, d.Nsw))) return rc;
    if ((rc = dalloc(c, s, &dcn, d.Nsw))) return rc;
    if ((rc = upload(c, dch, ch.data(), ch.size()))) return rc;
    if ((rc = upload(c, dct, ct.data(), ct.size()))) return rc;
    if ((rc = upload(c, dc0, c0.data(), d.Nsw))) return rc;
    if ((rc = upload(c, dcn, cn.data(), d.Nsw))) return rc;
    s->chunks = dch; s->ch_tile = dct; s->sw_chunk0 = dc0; s->sw_nchunk = dcn;
    if ((rc = dalloc(c, s, &s->ch_kpart, 6 * ch.size()))) return rc;
    if ((rc = dalloc(c, s, &s->ch_zpart, 6 * ch.size()))) return rc;
  }
  // Gauss-Jordan in LDS when one workgroup can hold the filament matrix
  {
    int lds_max = 0;
    if ((rc = hip_check(c, hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock,
                                                 c->device), "LDS size")))
      return rc;
    const size_t need = gj_lds_bytes(d.S) + 64;  // + the kernel's static shared scalars
    s->gj_lds = need <= size_t(lds_max);
    if (s->gj_lds && need > 65536 &&
        hipFuncSetAttribute(reinterpret_cast<const void *>(k_gj_invert_lds),
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            int(gj_lds_bytes(d.S))) != hipSuccess) {
      (void)hipGetLastError();
      s->gj_lds = false;
    }
  }
  if ((rc = dalloc(c, s, &s->tpart, 6 * std::max<size_t>(bt.size(), 1)))) return rc;
  if ((rc = dalloc(c, s, &s->zpart, 6 * std::max<size_t>(bt.size(), 1)))) return rc;
  // k_kprod reads the blob-blob work list's tile i as blob tile i
  {
    const WorkList &wl = c->lists[kBB];
    bool same = wl.ntiles == int(bt.size());
    for (int i = 0; same && i < wl.ntiles; i++)
      same = wl.h_tiles[i].t0 == bt[i].m0 && wl.h_tiles[i].nt == bt[i].n;
    if (!same) return fail(c, GMI_ERR_STATE, "blob tiles differ from the blob-blob work list");
  }
  return hip_check(c, hipStreamSynchronize(c->stream), "sync");
}

int gmi_sim_set_graphs(gmi_ctx *c, int use_graphs, int chunk) {
  SimHost *s = c ? sim_of(c) : nullptr;
  if (!s) return c ? fail(c, GMI_ERR_STATE, "no sim") : GMI_ERR_ARG;
  if (chunk < 1 || chunk > 64) return fail(c, GMI_ERR_ARG, "chunk must be in [1, 64]");
  int rc;
  if ((rc = hip_check(c, hipStreamSynchronize(c->stream), "sync"))) return rc;
  s->drop_graphs();
  s->use_graphs = use_graphs != 0;
  s->chunk = chunk;
  return GMI_OK;
}

int gmi_sim_graph_stats(gmi_ctx *c, int *captured_iterations, int *graphs) {
  SimHost *s = c ? sim_of(c) : nullptr;
  if (!s) return c ? fail(c, GMI_ERR_STATE, "no sim") : GMI_ERR_ARG;
  if (!captured_iterations || !graphs) return fail(c, GMI_ERR_ARG, "null output");
  *captured_iterations = s->captured_iters;
  *graphs = int(s->graphs.size() + s->outer_graphs.size());
  return GMI_OK;
}

int gmi_sim_set_body_preconditioner(gmi_ctx *c, int kind) {
  SimHost *s = c ? sim_of(c) : nullptr;
  if (!s) return c ? fail(c, GMI_ERR_STATE, "no sim") : GMI_ERR_ARG;
  if (kind != GMI_PRECOND_SELF_MOBILITY && kind != GMI_PRECOND_BODY_INVERSE)
    return fail(c, GMI_ERR_ARG, "unknown preconditioner");
  int rc;
  if ((rc = hip_check(c, hipStreamSynchronize(c->stream), "sync"))) return rc;
  s->drop_graphs();  // captured chunks hold the old preconditioner's kernels
  dense_release(s);
  // the exact inverse converges in about one Arnoldi step: check after every
  // step rather than launch (flag-skipped) steps that are not needed
  s->chunk = kind == GMI_PRECOND_BODY_INVERSE ? 1 : 6;
  if (kind == GMI_PRECOND_BODY_INVERSE) return dense_setup(c, s);
  return GMI_OK;
}

int gmi_sim_set_jacobian_pivoting(gmi_ctx *c, int always_pivot) {
  SimHost *s = c ? sim_of(c) : nullptr;
  if (!s) return c ? fail(c, GMI_ERR_STATE, "no sim") : GMI_ERR_ARG;
  if (always_pivot != 0 && always_pivot != 1) return fail(c, GMI_ERR_ARG, "always_pivot is 0 or 1");
  s->gj_pivot_always = always_pivot != 0;
  return GMI_OK;
}

int gmi_sim_set_gram_schmidt(gmi_ctx *c, int passes) {
  SimHost *s = c ? sim_of(c) : nullptr;
  if (!s) return c ? fail(c, GMI_ERR_STATE, "no sim") : GMI_ERR_ARG;
  if (passes != 1 && passes != 2) return fail(c, GMI_ERR_ARG, "passes is 1 or 2");
  int rc;
  if ((rc = hip_check(c, hipStreamSynchronize(c->stream), "sync"))) return rc;
  s->drop_graphs();  // captured chunks hold the other pass count's kernels
  s->gs_passes = passes;
  return GMI_OK;
}

int gmi_sim_set_gmres_history(gmi_ctx *c, int m) {
  SimHost *s = c ? sim_of(c) : nullptr;
  if (!s) return c ? fail(c, GMI_ERR_STATE, "no sim") : GMI_ERR_ARG;
  if (m < 0 || m > kHistMax) return fail(c, GMI_ERR_ARG, "history length must be in [0, 16]");
  int rc;
  if ((rc = hip_check(c, hipStreamSynchronize(c->stream), "sync"))) return rc;
  if (m > 0 && !s->hist_b) {
    const size_t n = size_t(s->d.npad), M = kHistMax, Nsw = size_t(s->d.Nsw);
    if ((rc = dalloc(c, s, &s->hist_b, M * n)) || (rc = dalloc(c, s, &s->hist_x, M * n)) ||
        (rc = dalloc(c, s, &s->hist_bnew, n)) || (rc = dalloc(c, s, &s->hist_G, Nsw * M * M)) ||
        (rc = dalloc(c, s, &s->hist_Gw, Nsw * M * M)) ||
        (rc = dalloc(c, s, &s->hist_part, size_t(s->nbtiles) * (M + 1))) ||
        (rc = dalloc(c, s, &s->hist_coef, Nsw * M)) ||
        (rc = dalloc(c, s, &s->hist_gsave, Nsw * (M + 1))) || (rc = dalloc(c, s, &s->hist_st, 2)))
      return rc;
  }
  // a new length starts an empty history; captured iterations hold the old
  // length's kernels
  s->drop_graphs();
  s->hist_m = m;
  if (s->hist_st && (rc = hip_check(c, hipMemsetAsync(s->hist_st, 0, 2 * sizeof(int), c->stream),
                                    "history reset")))
    return rc;
  return hip_check(c, hipStreamSynchronize(c->stream), "sync");
}

int gmi_sim_set_precision(gmi_ctx *c, int precision) {
  if (!c || !sim_of(c)) return GMI_ERR_STATE;
  if (precision != GMI_FP64 && precision != GMI_FP32_KAHAN)
    return fail(c, GMI_ERR_ARG, "unknown precision");
  SimHost *s = sim_of(c);
  if (s->precision != precision) {
    // captured Arnoldi chunks hold the kernels chosen for the old precision
    int rc;
    if ((rc = hip_check(c, hipStreamSynchronize(c->stream), "sync"))) return rc;
    s->drop_graphs();
  }
  s->precision = precision;
  return GMI_OK;
}

int gmi_make_body_reference_matrices(gmi_ctx *c) {
  SimHost *s = c ? sim_of(c) : nullptr;
  if (!s) return c ? fail(c, GMI_ERR_STATE, "no sim") : GMI_ERR_ARG;
  SimDev &d = s->d;
  int rc;
  // (K^T K)^-1 per swimmer from its body-frame blob references (host, 6x6)
  std::vector<double> body(size_t(d.Nsw) * BREC);
  if ((rc = hip_check(c, hipStreamSynchronize(c->stream), "sync")) ||
      (rc = hip_check(c, hipMemcpy(body.data(), d.body, sizeof(double) * body.size(),
                                   hipMemcpyDeviceToHost), "body D2H")))
    return rc;
  for (int n = 0; n < d.Nsw; n++) {
    double K[36] = {0};
    const int nb = c->nblob[n];
    K[0] = K[7] = K[14] = nb;
    for (int m = 0; m < nb; m++) {
      const double *r = s->p.blob_refs + 3 * (c->blob_off[n] + m);
      // rcross(r) = [[0, r2, -r1], [-r2, 0, r0], [r1, -r0, 0]]
      const double rc3[9] = {0, r[2], -r[1], -r[2], 0, r[0], r[1], -r[0], 0};
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
          K[6 * i + 3 + j] += rc3[3 * i + j];
          K[6 * (3 + i) + j] -= rc3[3 * i + j];
          double sq = 0;
          for (int k = 0; k < 3; k++) sq += rc3[3 * i + k] * rc3[3 * k + j];
          K[6 * (3 + i) + 3 + j] -= sq;
        }
    }
    // host Gauss-Jordan, partial pivoting
    int piv[6] = {0, 1, 2, 3, 4, 5};
    for (int k = 0; k < 6; k++) {
      int pp = k;
      double best = std::fabs(K[6 * k + k]);
      for (int i = k + 1; i < 6; i++)
        if (std::fabs(K[6 * i + k]) > best) { best = std::fabs(K[6 * i + k]); pp = i; }
      if (best == 0.0) return fail(c, GMI_ERR_SINGULAR, "K^T K is singular");
      if (pp != k) {
        for (int j = 0; j < 6; j++) std::swap(K[6 * k + j], K[6 * pp + j]);
        std::swap(piv[k], piv[pp]);
      }
      const double inv = 1.0 / K[6 * k + k];
      K[6 * k + k] = 1.0;
      for (int j = 0; j < 6; j++) K[6 * k + j] *= inv;
      for (int i = 0; i < 6; i++) {
        if (i == k) continue;
        const double fct = K[6 * i + k];
        if (fct == 0.0) continue;
        K[6 * i + k] = 0.0;
        for (int j = 0; j < 6; j++) K[6 * i + j] -= fct * K[6 * k + j];
      }
    }
    double* b = &body[size_t(n) * BREC];
    for (int j = 0; j < 6; j++)
      for (int i = 0; i < 6; i++) b[BK + 6 * i + piv[j]] = K[6 * i + j];
  }
  if ((rc = upload(c, d.body, body.data(), body.size()))) return rc;
  if ((rc = read_positions_and_forces(c, s))) return rc;
  for (int n = 0; n < 6; n++) {
    hipLaunchKernelGGL(k_unit_rhs, dim3(nblk(d.npad, 256)), dim3(256), 0, c->stream, d, s->rhs,
                       5 - n);
    if ((rc = launch_ok(c, "unit rhs"))) return rc;
    int it = 0;
    if ((rc = solve_linear_system(c, s, &it))) return rc;
    hipLaunchKernelGGL(k_store_nref_col, dim3(nblk(d.Nsw, 64)), dim3(64), 0, c->stream, d, s->x,
                       5 - n);
    if ((rc = launch_ok(c, "store N_ref"))) return rc;
  }
  return hip_check(c, hipStreamSynchronize(c->stream), "sync");
}

int gmi_initial_guess(gmi_ctx *c, int nt) {
  SimHost *s = c ? sim_of(c) : nullptr;
  if (!s) return c ? fail(c, GMI_ERR_STATE, "no sim") : GMI_ERR_ARG;
  hipLaunchKernelGGL(k_body_initial_guess, dim3(nblk(s->d.Nsw, 64)), dim3(64), 0, c->stream, s->d,
                     nt);
  hipLaunchKernelGGL(k_fil_initial_guess, dim3(s->d.Nf), dim3(kFilLanes), 0, c->stream, s->d,
                     nt);
  return launch_ok(c, "initial guess");
}

// forces at the step set on the device (set_step)
static int forces_and_torques(gmi_ctx *c, SimHost *s, int perturb) {
  hipLaunchKernelGGL(k_fil_forces, dim3(s->d.Nf), dim3(kFilLanes), 0, c->stream, s->d, perturb);
  hipLaunchKernelGGL(k_body_forces, dim3(s->d.Nsw), dim3(256), 0, c->stream, s->d);
  return launch_ok(c, "forces");
}

int gmi_forces_and_torques(gmi_ctx *c, int nt) {
  SimHost *s = c ? sim_of(c) : nullptr;
  if (!s) return c ? fail(c, GMI_ERR_STATE, "no sim") : GMI_ERR_ARG;
  int rc;
  if ((rc = set_step(c, s, nt))) return rc;
  return forces_and_torques(c, s, (nt == 0 && s->d.pert) ? 1 : 0);
}

int gmi_compute_velocities(gmi_ctx *c, int nt, int *gmres_iters) {
  (void)nt;
  SimHost *s = c ? sim_of(c) : nullptr;
  if (!s) return c ? fail(c, GMI_ERR_STATE, "no sim") : GMI_ERR_ARG;
  int rc = compute_velocities(c, s, gmres_iters);
  if (rc) return rc;
  return hip_check(c, hipStreamSynchronize(c->stream), "sync");
}

int gmi_compute_errors(gmi_ctx *c, int nt, int which, int *too_large, int *not_finite) {
  SimHost *s = c ? sim_of(c) : nullptr;
  if (!s) return c ? fail(c, GMI_ERR_STATE, "no sim") : GMI_ERR_ARG;
  if (which != 0 && which != 1) return fail(c, GMI_ERR_ARG, "which must be 0 or 1");
  int rc;
  if ((rc = set_step(c, s, nt))) return rc;
  return compute_errors(c, s, which ? s->nerr : s->err, too_large, not_finite);
}

int gmi_prepare_jacobian_inv(gmi_ctx *c, int nt) {
  SimHost *s = c ? sim_of(c) : nullptr;
  if (!s) return c ? fail(c, GMI_ERR_STATE, "no sim") : GMI_ERR_ARG;
  SimDev &d = s->d;
  const double t = (nt + 1) * s->p.dt;
  hipLaunchKernelGGL(k_set_ints, dim3(1), dim3(64), 0, c->stream, d.flags + F_SING, 1, 0,
                     nullptr);
  hipLaunchKernelGGL(k_fil_jacobian, dim3(d.Nf), dim3(64), 0, c->stream, d, nt, t);
  if (6 * d.S <= kGjRegMax && s->gj_pivot_always) {
    hipLaunchKernelGGL(k_gj_invert_reg, dim3(d.Nf), dim3(256), 0, c->stream, d, nullptr);
  } else if (6 * d.S <= kGjRegMax) {
    // pivot-free first; partial pivoting for the filaments whose inverse
    // fails the check
    hipLaunchKernelGGL(k_gj_nopiv_reg, dim3(d.Nf), dim3(256), 0, c->stream, d);
    hipLaunchKernelGGL(k_gj_invert_reg, dim3(d.Nf), dim3(256), 0, c->stream, d, d.gj_redo);
  }
  else if (s->gj_lds)
    hipLaunchKernelGGL(k_gj_invert_lds, dim3(d.Nf), dim3(kGjThreads), gj_lds_bytes(d.S), c->stream,
                       d);
  else
    hipLaunchKernelGGL(k_gj_invert, dim3(d.Nf), dim3(256), 0, c->stream, d);
  hipLaunchKernelGGL(k_schur_fil, dim3(d.Nf), dim3(256), 0, c->stream, d, nt);
  hipLaunchKernelGGL(k_schur_body, dim3(d.Nsw), dim3(256), 0, c->stream, d, nt);
  int rc;
  if ((rc = launch_ok(c, "prepare_jacobian_inv"))) return rc;
  if ((rc = read_flags(c, s))) return rc;
  if (s->h_flags[F_SING]) return fail(c, GMI_ERR_SINGULAR, "singular approximate Jacobian");
  s->iter = 0;
  return GMI_OK;
}

// -H_k e: -gamma J^-1 e - sum_i (D_i . e) C_i (broyden_solver.cpp:26-67)
static int find_update(gmi_ctx *c, SimHost *s) {
  SimDev &d = s->d;
  int rc;
  if ((rc = jacobian_inv_mult(c, s, s->err, s->upd, nullptr))) return rc;
  const int n = d.nbroy;
  if (s->iter > 0) {
    const DotShape ds = dot_shape(n);
    hipLaunchKernelGGL(k_pdots, dim3(ds.nb), dim3(256), sizeof(double) * ds.R, c->stream, s->D,
                       size_t(n), s->iter, s->err, n, ds.R, s->pp, nullptr);
    if (ds.wide)
      hipLaunchKernelGGL(k_colsum_wide, dim3(s->iter), dim3(256), 0, c->stream, s->pp, ds.nb,
                         s->bd, nullptr);
    else
      hipLaunchKernelGGL(k_colsum, dim3(nblk(s->iter, 256)), dim3(256), 0, c->stream, s->pp,
                         s->iter, ds.nb, s->bd, nullptr);
  }
  hipLaunchKernelGGL(k_broyden_combine, dim3(nblk(n, 256)), dim3(256), 0, c->stream, s->upd,
                     s->upd, s->p.jacobian_confidence, s->C, size_t(n), s->bd, s->iter, n,
                     nullptr);
  return launch_ok(c, "find_update");
}

int gmi_broyden_find_update(gmi_ctx *c) {
  SimHost *s = c ? sim_of(c) : nullptr;
  if (!s) return c ? fail(c, GMI_ERR_STATE, "no sim") : GMI_ERR_ARG;
  return find_update(c, s);
}

static int apply_update(gmi_ctx *c, SimHost *s) {
  hipLaunchKernelGGL(k_body_update, dim3(nblk(s->d.Nsw, 64)), dim3(64), 0, c->stream, s->d, s->upd);
  hipLaunchKernelGGL(k_fil_update, dim3(s->d.Nf), dim3(kFilLanes), 0, c->stream, s->d, s->upd);
  return launch_ok(c, "update");
}

int gmi_apply_update(gmi_ctx *c) {
  SimHost *s = c ? sim_of(c) : nullptr;
  if (!s) return c ? fail(c, GMI_ERR_STATE, "no sim") : GMI_ERR_ARG;
  return apply_update(c, s);
}

// The new Broyden pair C_k, D_k from err (e) and nerr (e_new)
// (broyden_solver.cpp:70-130): C_k = -H_k e_new / |e_new - e|,
// D_k = (e_new - e) / |e_new - e|; skip gates every kernel.
static int enqueue_end_of_iter(gmi_ctx *c, SimHost *s, const int *skip) {
  SimDev &d = s->d;
  const int n = d.nbroy;
  int rc;
  if ((rc = jacobian_inv_mult(c, s, s->nerr, s->upd, skip))) return rc;
  const int it = s->iter;
  double *Cit = s->C + size_t(n) * it, *Dit = s->D + size_t(n) * it;
  if (it > 0) {
    const DotShape ds = dot_shape(n);
    // k_pdots and k_colsum_wide return when their flag word is set
    hipLaunchKernelGGL(k_pdots, dim3(ds.nb), dim3(256), sizeof(double) * ds.R, c->stream, s->D,
                       size_t(n), it, s->nerr, n, ds.R, s->pp, skip);
    if (ds.wide)
      hipLaunchKernelGGL(k_colsum_wide, dim3(it), dim3(256), 0, c->stream, s->pp, ds.nb, s->bd,
                         skip);
    else
      hipLaunchKernelGGL(k_colsum, dim3(nblk(it, 256)), dim3(256), 0, c->stream, s->pp, it,
                         ds.nb, s->bd, skip);
  }
  hipLaunchKernelGGL(k_broyden_combine, dim3(nblk(n, 256)), dim3(256), 0, c->stream, Cit, s->upd,
                     s->p.jacobian_confidence, s->C, size_t(n), s->bd, it, n, skip);
  hipLaunchKernelGGL(k_axpby, dim3(nblk(n, 256)), dim3(256), 0, c->stream, Dit, 1.0, s->nerr,
                     -1.0, s->err, n, skip);
  hipLaunchKernelGGL(k_norms, dim3(1), dim3(1024), 0, c->stream, Dit, nullptr, n, s->scal, 8, 9,
                     skip);
  hipLaunchKernelGGL(k_scale2, dim3(nblk(n, 256)), dim3(256), 0, c->stream, Cit, Dit, n, s->scal,
                     8, skip);
  return launch_ok(c, "end_of_iter");
}

// host side of end_of_iter: the new pair is kept when the residual was too
// large and there is room for it
static bool stores_pair(const SimHost *s, int too_large) {
  return too_large && s->iter < s->p.max_broyden_iter - 1;
}

int gmi_broyden_end_of_iter(gmi_ctx *c, int too_large) {
  SimHost *s = c ? sim_of(c) : nullptr;
  if (!s) return c ? fail(c, GMI_ERR_STATE, "no sim") : GMI_ERR_ARG;
  if (stores_pair(s, too_large)) {
    int rc;
    if ((rc = enqueue_end_of_iter(c, s, nullptr))) return rc;
    std::swap(s->err, s->nerr);
  }
  s->iter++;
  return GMI_OK;
}

// ---- captured Broyden iterations ------------------------------------------
// One graph per (Broyden index, which buffer err is, precision, preconditioner)
// holds a whole iteration of cilia_sim_main.cpp:266-314: find_update, the
// state update, the forces, the velocities up to the first GMRES chunk, and,
// gated on that chunk having converged (F_HOLD), the rest of the solve, the
// segment velocities, the residual and the new Broyden pair (gated on the
// residual being finite and too large).  The host reads the flags once per
// iteration; a solve that needs more than the first chunk is finished with
// eager launches of the same kernels.  The step's time and index are read on
// the device, so the graphs serve every step; so are the fitted guess's
// history count and head.  Captured when GMRES runs in FP64 (plain FP64, or
// under the exact body-frame preconditioner), from the second step on (the
// first step's forces carry the nt = 0 perturbations).
static bool iteration_capturable(gmi_ctx *c, SimHost *s, int nt) {
  return s->use_graphs && !c->profiling && !(nt == 0 && s->d.pert) &&
         (s->precision == GMI_FP64 || s->dense != nullptr);
}

static int gmres_maxit(const SimHost *s) {
  return std::min(s->d.maxit, 3 * s->d.Nb + 6 * s->d.Nsw);
}

static int enqueue_iteration(gmi_ctx *c, SimHost *s) {
  SimDev &d = s->d;
  int rc;
  if ((rc = find_update(c, s))) return rc;
  if ((rc = apply_update(c, s))) return rc;
  if ((rc = forces_and_torques(c, s, 0))) return rc;
  if ((rc = velocities_before_solve(c, s))) return rc;
  hipLaunchKernelGGL(k_rotate_body, dim3(nblk(d.Nsw, 64)), dim3(64), 0, c->stream, d, s->rot);
  if ((rc = hist_guess(c, s))) return rc;
  SimOp op{c, s};
  const int maxit = gmres_maxit(s);
  const double tol = s->p.gmres_tol;
  if ((rc = gmres_begin(c, s, op, d.npad, tol))) return rc;
  if ((rc = gmres_enqueue(c, s, op, d.npad, maxit, false, tol, 1, std::min(maxit, s->chunk))))
    return rc;
  hipLaunchKernelGGL(k_hold_unsolved, dim3(1), dim3(1), 0, c->stream, d.flags);
  const int *hold = d.flags + F_HOLD;
  if ((rc = gmres_finish(c, s, op, d.npad, hold))) return rc;
  if ((rc = hist_push(c, s, hold))) return rc;
  if ((rc = velocities_after_solve(c, s, hold))) return rc;
  if ((rc = enqueue_errors(c, s, s->nerr, hold))) return rc;
  if (s->iter < s->p.max_broyden_iter - 1) {
    hipLaunchKernelGGL(k_hold_unless_big, dim3(1), dim3(1), 0, c->stream, d.flags);
    if ((rc = enqueue_end_of_iter(c, s, hold))) return rc;
  }
  return launch_ok(c, "captured iteration");
}

// One Broyden iteration from a graph: *big, *bad from the residual, *its the
// GMRES iterations.  Leaves the host side of end_of_iter to the caller.
static int captured_iteration(gmi_ctx *c, SimHost *s, int *big, int *bad, int *its) {
  int rc;
  if (s->graph_gen != c->list_generation) {
    s->drop_graphs();
    s->graph_gen = c->list_generation;
  }
  const long long key = (static_cast<long long>(s->iter) << 3) | (s->err == s->err_first ? 1 : 0) |
                        (s->precision == GMI_FP32_KAHAN ? 2 : 0) | (s->dense ? 4 : 0);
  auto g = s->outer_graphs.find(key);
  if (g == s->outer_graphs.end()) {
    CapturedGraph cg;
    if ((rc = capture_graph(c, [&]() { return enqueue_iteration(c, s); }, &cg))) return rc;
    g = s->outer_graphs.emplace(key, cg).first;
  }
  if ((rc = hip_check(c, hipGraphLaunch(g->second.exec, c->stream), "graph launch"))) return rc;
  s->captured_iters++;
  if ((rc = read_flags(c, s))) return rc;
  if (s->h_flags[F_DONE]) {
    *its = s->h_flags[F_ITS];
    *big = s->h_flags[F_BIG];
    *bad = s->h_flags[F_NAN];
  } else {
    // the solve needs more than the first chunk: the rest of it and the
    // kernels the graph held back, eagerly
    SimOp op{c, s};
    const int maxit = gmres_maxit(s);
    const double tol = s->p.gmres_tol;
    if ((rc = gmres_chunks(c, s, op, s->d.npad, maxit, false, tol, std::min(maxit, s->chunk) + 1,
                           its)))
      return rc;
    if ((rc = gmres_finish(c, s, op, s->d.npad, nullptr))) return rc;
    if ((rc = hist_push(c, s, nullptr))) return rc;
    if ((rc = velocities_after_solve(c, s, nullptr))) return rc;
    if ((rc = compute_errors(c, s, s->nerr, big, bad))) return rc;
    if (!*bad && stores_pair(s, *big) && (rc = enqueue_end_of_iter(c, s, nullptr))) return rc;
  }
  s->last_gmres = *its;
  return GMI_OK;
}

// Broyden iterations from the current iterate (cilia_sim_main.cpp:252-314):
// returns GMI_OK, GMI_ERR_NOT_CONVERGED after max_iter iterations, or
// GMI_ERR_NOT_FINITE; sets no error message for those two.  Like the
// reference driver, a non-finite first residual is only caught when
// `check_first` is set (the learned attempt).
static int broyden_solve(gmi_ctx *c, SimHost *s, int nt, int max_iter, bool check_first,
                         int *iters, int *gmres_total) {
  int rc, big = 0, bad = 0, its = 0;
  if ((rc = gmi_forces_and_torques(c, nt))) return rc;
  if ((rc = compute_velocities(c, s, &its))) return rc;
  *gmres_total += its;
  if ((rc = compute_errors(c, s, s->err, &big, &bad))) return rc;
  *iters = 0;
  if (bad && check_first) return GMI_ERR_NOT_FINITE;
  if (big && (rc = gmi_prepare_jacobian_inv(c, nt))) return rc;
  s->iter = 0;
  const bool captured = iteration_capturable(c, s, nt);
  const int perturb = (nt == 0 && s->d.pert) ? 1 : 0;
  while (big && s->iter < max_iter) {
    if (captured) {
      if ((rc = captured_iteration(c, s, &big, &bad, &its))) return rc;
    } else {
      if ((rc = find_update(c, s))) return rc;
      if ((rc = apply_update(c, s))) return rc;
      if ((rc = forces_and_torques(c, s, perturb))) return rc;
      if ((rc = compute_velocities(c, s, &its))) return rc;
      if ((rc = compute_errors(c, s, s->nerr, &big, &bad))) return rc;
      if (!bad && stores_pair(s, big) && (rc = enqueue_end_of_iter(c, s, nullptr))) return rc;
    }
    *gmres_total += its;
    if (bad) {
      *iters = s->iter + 1;
      return GMI_ERR_NOT_FINITE;
    }
    // host side of end_of_iter
    if (stores_pair(s, big)) std::swap(s->err, s->nerr);
    s->iter++;
  }
  *iters = s->iter;
  return big ? GMI_ERR_NOT_CONVERGED : GMI_OK;
}

static void apply_update_vector(gmi_ctx *c, SimHost *s, const double *v) {
  hipLaunchKernelGGL(k_body_update, dim3(nblk(s->d.Nsw, 64)), dim3(64), 0, c->stream, s->d, v);
  hipLaunchKernelGGL(k_fil_update, dim3(s->d.Nf), dim3(kFilLanes), 0, c->stream, s->d, v);
}

int gmi_step(gmi_ctx *c, int nt, int *broyden_iters, int *gmres_iters) {
  SimHost *s = c ? sim_of(c) : nullptr;
  if (!s) return c ? fail(c, GMI_ERR_STATE, "no sim") : GMI_ERR_ARG;
  int rc, bi = 0, total = 0, spent = 0;
  if ((rc = gmi_initial_guess(c, nt))) return rc;
  bool corrected = false;
  if ((rc = learn_after_initial_guess(c, s, nt, &corrected))) return rc;
  rc = GMI_ERR_NOT_CONVERGED;
  if (corrected) {
    // network proposal: the extrapolated iterate plus the predicted
    // correction, applied through the ordinary state update.  The attempt
    // gets about twice the recent iteration count; if it does not
    // converge in that, the step restarts from the extrapolated guess.
    apply_update_vector(c, s, learn_correction(s));
    if ((rc = launch_ok(c, "warm start"))) return rc;
    const int cap = std::min(s->p.max_broyden_iter, std::max(4, int(2.0 * s->iters_avg) + 2));
    rc = broyden_solve(c, s, nt, cap, true, &bi, &total);
    spent = bi;
    if (rc != GMI_OK && rc != GMI_ERR_NOT_CONVERGED && rc != GMI_ERR_NOT_FINITE) return rc;
    if (rc != GMI_OK) {
      double *back = nullptr;
      if ((rc = learn_back_to_guess(c, s, &back))) return rc;
      apply_update_vector(c, s, back);
      if ((rc = launch_ok(c, "back to guess"))) return rc;
      rc = GMI_ERR_NOT_CONVERGED;
    }
  }
  if (rc != GMI_OK) {
    rc = broyden_solve(c, s, nt, s->p.max_broyden_iter, false, &bi, &total);
    spent += bi;
  }
  if (broyden_iters) *broyden_iters = spent;
  if (gmres_iters) *gmres_iters = total;
  if (rc == GMI_ERR_NOT_FINITE) return fail(c, rc, "Broyden's method diverged");
  if (rc == GMI_ERR_NOT_CONVERGED) return fail(c, rc, "Broyden's method did not converge");
  if (rc) return rc;
  s->iters_avg = 0.8 * s->iters_avg + 0.2 * spent;
  return learn_after_convergence(c, s, nt);
}

int gmi_get_state(gmi_ctx *c, double *seg_x, double *seg_q, double *seg_u, double *fil_lam,
                  double *body_xqu) {
  SimHost *s = c ? sim_of(c) : nullptr;
  if (!s) return c ? fail(c, GMI_ERR_STATE, "no sim") : GMI_ERR_ARG;
  SimDev &d = s->d;
  int rc;
  if ((rc = hip_check(c, hipStreamSynchronize(c->stream), "sync"))) return rc;
  const size_t ls = size_t(d.lds);
  std::vector<double> buf;
  auto soa = [&](const double *src, int nc, double *dst) -> int {
    buf.resize(nc * ls);
    int r = hip_check(c, hipMemcpy(buf.data(), src, sizeof(double) * nc * ls, hipMemcpyDeviceToHost),
                      "state D2H");
    if (r) return r;
    for (int g = 0; g < d.Ns; g++)
      for (int k = 0; k < nc; k++) dst[nc * g + k] = buf[k * ls + g];
    return GMI_OK;
  };
  if (seg_x && (rc = soa(d.X, 3, seg_x))) return rc;
  if (seg_q && (rc = soa(d.Q, 4, seg_q))) return rc;
  if (seg_u && (rc = soa(d.U, 3, seg_u))) return rc;
  if (fil_lam) {
    const int S = d.S;
    std::vector<double> L(size_t(d.Nf) * 9 * S), T(9 * size_t(d.Nf)), Cl(9 * size_t(d.Nf));
    if ((rc = hip_check(c, hipMemcpy(L.data(), d.lam, sizeof(double) * L.size(), hipMemcpyDeviceToHost), "D2H")))
      return rc;
    if ((rc = hip_check(c, hipMemcpy(T.data(), d.teth, sizeof(double) * T.size(), hipMemcpyDeviceToHost), "D2H")))
      return rc;
    if ((rc = hip_check(c, hipMemcpy(Cl.data(), d.clamp, sizeof(double) * Cl.size(), hipMemcpyDeviceToHost), "D2H")))
      return rc;
    for (int f = 0; f < d.Nf; f++) {
      double *o = fil_lam + size_t(f) * (3 * S + 3);
      for (int k = 0; k < 3 * (S - 1); k++) o[k] = L[size_t(f) * 9 * S + k];
      for (int k = 0; k < 3; k++) { o[3 * (S - 1) + k] = T[9 * f + k]; o[3 * S + k] = Cl[9 * f + k]; }
    }
  }
  if (body_xqu) {
    std::vector<double> B(size_t(d.Nsw) * BREC);
    if ((rc = hip_check(c, hipMemcpy(B.data(), d.body, sizeof(double) * B.size(), hipMemcpyDeviceToHost), "D2H")))
      return rc;
    for (int n = 0; n < d.Nsw; n++) {
      const double *b = &B[size_t(n) * BREC];
      double *o = body_xqu + 10 * n;
      for (int k = 0; k < 3; k++) o[k] = b[BX + k];
      for (int k = 0; k < 4; k++) o[3 + k] = b[BQ + k];
      for (int k = 0; k < 3; k++) o[7 + k] = b[BU + k];
    }
  }
  return GMI_OK;
}

int gmi_get_mobility(gmi_ctx *c, double *v_segs, double *v_bodies, double *f_blobs,
                     double *body_ref66) {
  SimHost *s = c ? sim_of(c) : nullptr;
  if (!s) return c ? fail(c, GMI_ERR_STATE, "no sim") : GMI_ERR_ARG;
  SimDev &d = s->d;
  int rc;
  if ((rc = hip_check(c, hipStreamSynchronize(c->stream), "sync"))) return rc;
  if (v_segs) {
    std::vector<double> buf(6 * size_t(d.lds));
    if ((rc = hip_check(c, hipMemcpy(buf.data(), c->Vs, sizeof(double) * buf.size(), hipMemcpyDeviceToHost), "D2H")))
      return rc;
    for (int g = 0; g < d.Ns; g++)
      for (int k = 0; k < 6; k++) v_segs[6 * g + k] = buf[size_t(k) * d.lds + g];
  }
  std::vector<double> xv(d.npad);
  if ((rc = hip_check(c, hipMemcpy(xv.data(), s->x, sizeof(double) * d.npad, hipMemcpyDeviceToHost), "D2H")))
    return rc;
  if (v_bodies)
    for (int k = 0; k < 6 * d.Nsw; k++) v_bodies[k] = xv[3 * d.ldb + k];
  if (f_blobs)
    for (int m = 0; m < d.Nb; m++)
      for (int k = 0; k < 3; k++) f_blobs[3 * m + k] = xv[size_t(k) * d.ldb + m];
  if (body_ref66) {
    std::vector<double> B(size_t(d.Nsw) * BREC);
    if ((rc = hip_check(c, hipMemcpy(B.data(), d.body, sizeof(double) * B.size(), hipMemcpyDeviceToHost), "D2H")))
      return rc;
    for (int n = 0; n < d.Nsw; n++)
      for (int k = 0; k < 36; k++) {
        body_ref66[72 * n + k] = B[size_t(n) * BREC + BN + k];
        body_ref66[72 * n + 36 + k] = B[size_t(n) * BREC + BK + k];
      }
  }
  return GMI_OK;
}

int gmi_get_broyden(gmi_ctx *c, int which, double *out) {
  SimHost *s = c ? sim_of(c) : nullptr;
  if (!s) return c ? fail(c, GMI_ERR_STATE, "no sim") : GMI_ERR_ARG;
  if (!out || which < 0 || which > 2) return fail(c, GMI_ERR_ARG, "bad arguments");
  const double *src = which == 0 ? s->err : (which == 1 ? s->nerr : s->upd);
  // a null-stream copy does not wait for the context's (non-blocking) stream
  int rc;
  if ((rc = hip_check(c, hipStreamSynchronize(c->stream), "sync"))) return rc;
  return hip_check(c, hipMemcpy(out, src, sizeof(double) * s->d.nbroy, hipMemcpyDeviceToHost),
                   "D2H");
}

}  // extern "C"
