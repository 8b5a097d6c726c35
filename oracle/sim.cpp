// sim.cpp -- oracle (test infrastructure only): FP64 CPU restatement of the
// reference's implicit time step for free-swimming ciliated spheres:
//   driver loop            cilia_sim_main.cpp:205-322
//   Broyden (inverse form) broyden_solver.cpp:26-130
//   swimmer                swimmer.cpp:240-325, 329-688 (GMRES branch 466-682),
//                          690-845 (793-839), 847-871
//   filament               filament.cpp:8-55, 416-463, 825-1018, 1096-1575
//   segment                segment.cpp:38-90
//   rigid body             rigid_body.cpp:247-381
//   mobility solver        mobility_solver.cpp:132-192, 361-421, 651-727,
//                          905-964, 970-1354, 1358-1402, 1404-1689, 1693-1771
// Physics (kept identical on the device):
//  * the reference's default instability-driven cilia (CILIA_TYPE 0): a
//    follower force -END_FORCE_MAGNITUDE t at each filament tip and, at
//    nt = 0, the perturbations on segments 4 and 12 (filament.cpp:1177-1202),
//    with their approximate-Jacobian and Schur-complement terms
//    (filament.cpp:1466-1470; swimmer.cpp:497-503,557-571,630-633).  The
//    perturbation is an input (the reference draws it from
//    std::random_device, filament.cpp:336-360); end force 0 switches all of
//    this off;
//  * driving: preferred curvature kappa0(t) = amp sin(2 pi t / period + phi_f)
//    in the binormal component of strain_twist (filament.cpp:1104-1106),
//    evaluated at the end-of-step time t = (nt+1) dt (amp 0: none);
//  * swimmers couple through the pair sums and the barrier (interact = 1, the
//    reference's NSWIM > 1 run) or are hydrodynamically independent
//    (interact = 0); either way each swimmer has its own blob/filament
//    references and its own body-mobility reference matrices;
//  * the blob part of the preconditioner is scaled once per blob (the
//    reference rescales all blobs once per swimmer, mobility_solver.cpp:719,
//    which only matters for NSWIM > 1 and never changes the solution).
#include "oracle_sim.h"
#include "oracle.h"
#include "lie.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

using namespace orc;

namespace {

struct Segment {
  double x[3], xm1[3], xm2[3], u[3] = {0, 0, 0}, um1[3] = {0, 0, 0};
  Quat q, qm1;
};

struct Filament {
  std::vector<Segment> seg;
  std::vector<double> lam, lamm1, lamm2;  // 3(S-1)
  double teth[3] = {0, 0, 0}, tethm1[3] = {0, 0, 0}, tethm2[3] = {0, 0, 0};
  double clamp[3] = {0, 0, 0}, clampm1[3] = {0, 0, 0}, clampm2[3] = {0, 0, 0};
  double st[3] = {0, 0, 0};               // strain_twist
  double pert[6] = {0, 0, 0, 0, 0, 0};    // perturbation1, perturbation2
  double phase = 0;
  Mat Jinv;                               // 6S x 6S
  M3 ecb1, ecb2;                          // elastic clamping blocks
  double *f = nullptr;                    // -> global f_segs (6S)
};

struct Swimmer {
  int nfil = 0, nblob = 0;
  int seg0 = 0, blob0 = 0;     // offsets in the global particle arrays
  int err0 = 0;                // offset in the Broyden vectors
  double R = 0;                // body radius of the Schur RPY correction
  double x[3] = {0, 0, 0}, xm1[3] = {0, 0, 0}, xm2[3] = {0, 0, 0};
  double u[3] = {0, 0, 0}, um1[3] = {0, 0, 0};
  Quat q, qm1;
  M3 Qinit = eye3();
  std::vector<double> blob_ref;  // 3 nblob
  std::vector<double> fil_ref;   // 3 nfil
  std::vector<Filament> fil;
  double f[6] = {0, 0, 0, 0, 0, 0};
  Mat schur_inv;                 // 6x6
  std::vector<Mat> Bblk;         // nfil x (6 x 6S)
  Mat Nref, KTKinv;              // 6x6 each
};

}  // namespace

struct orc_sim {
  orc_sim_params p;
  int S = 0;
  int Ns = 0, Nb = 0, nbroy = 0, nsys = 0;
  std::vector<Swimmer> sw;
  // mobility-solver host buffers
  std::vector<double> xs, fs, vs, xb, fb, vb, frep, fs_dev;
  std::vector<double> v_bodies;
  // GMRES storage
  std::vector<double> rhs;
  // Broyden
  Mat C, D;
  std::vector<double> err, new_err, upd;
  int iter = 0;
  int last_gmres = 0;
  long long total_gmres = 0;
  int nthreads = 1;
  int nt_cur = 0;           // time step of the forces (set_driving)
  double DL() const { return 2.2 * p.a_seg; }
};

namespace {

// ----------------------------------------------------------------- filament

// filament.cpp:416-436
void robot_arm(Filament &F, double DL) {
  const int S = int(F.seg.size());
  double t1[3], t2[3];
  tangent(F.seg[0].q, t1);
  for (int i = 1; i < S; i++) {
    tangent(F.seg[i].q, t2);
    for (int k = 0; k < 3; k++) F.seg[i].x[k] = F.seg[i - 1].x[k] + 0.5 * DL * (t1[k] + t2[k]);
    std::memcpy(t1, t2, sizeof t1);
  }
}

// segment.cpp:81-90
void seg_update(Segment &s, const double *du) {
  for (int k = 0; k < 3; k++) s.u[k] += du[k];
  s.q = qmul(lie_exp(s.u), s.qm1);
}

// filament.cpp:438-463
void accept_state_from_rigid_body(Filament &F, const double *x_in, const double *u_in, double DL) {
  for (int k = 0; k < 3; k++) F.seg[0].x[k] = x_in[k];
  const double du[3] = {u_in[0] - F.seg[0].u[0], u_in[1] - F.seg[0].u[1], u_in[2] - F.seg[0].u[2]};
  seg_update(F.seg[0], du);
  robot_arm(F, DL);
}

// segment.cpp:38-79
void seg_initial_guess(Segment &s, int nt, int neuler) {
  std::memcpy(s.xm2, s.xm1, sizeof s.xm2);
  std::memcpy(s.xm1, s.x, sizeof s.xm1);
  if (nt <= neuler) {
    std::memcpy(s.um1, s.u, sizeof s.um1);
  } else {
    for (int k = 0; k < 3; k++) {
      const double t = 2.0 * s.u[k] - s.um1[k];
      s.um1[k] = s.u[k];
      s.u[k] = t;
    }
  }
  s.qm1 = s.q;
  s.q = qmul(lie_exp(s.u), s.qm1);
}

void extrap(double *v, double *m1, double *m2, int n, int nt) {
  for (int k = 0; k < n; k++) {
    if (nt == 1) {
      m1[k] = v[k];
    } else if (nt == 2) {
      m2[k] = m1[k]; m1[k] = v[k]; v[k] = 2.0 * m1[k] - m2[k];
    } else if (nt > 2) {
      const double t = 3.0 * (v[k] - m1[k]) + m2[k];
      m2[k] = m1[k]; m1[k] = v[k]; v[k] = t;
    }
  }
}

// filament.cpp:825-1018 (non-prescribed branch)
void fil_initial_guess(Filament &F, int nt, const double *x_in, const double *u_in, int neuler,
                       double DL) {
  for (Segment &s : F.seg) seg_initial_guess(s, nt, neuler);
  accept_state_from_rigid_body(F, x_in, u_in, DL);
  const int S = int(F.seg.size());
  extrap(F.lam.data(), F.lamm1.data(), F.lamm2.data(), 3 * (S - 1), nt);
  extrap(F.teth, F.tethm1, F.tethm2, 3, nt);
  extrap(F.clamp, F.clampm1, F.clampm2, 3, nt);
}

// filament.cpp:1096-1110
void body_frame_moment(const Filament &F, int i, double DL, double KB, double KT, double *M) {
  const Quat qm = qconj(midpoint(F.seg[i].q, F.seg[i + 1].q));
  const Quat k = qscale(qmul(qm, qscale(qsub(F.seg[i + 1].q, F.seg[i].q), 1.0 / DL)), 2.0);
  M[0] = KT * (k.v[0] - F.st[0]);
  M[1] = KB * (k.v[1] - F.st[1]);
  M[2] = KB * (k.v[2] - F.st[2]);
}

// psi_mat (quaternion.cpp:315-329), 4x3 as Mat
Mat psi_mat(const Quat &q) {
  Mat P(4, 3);
  P(0, 0) = -q.v[0]; P(0, 1) = -q.v[1]; P(0, 2) = -q.v[2];
  const M3 rc = rcross(q.v);
  for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) P(1 + i, j) = rc(i, j);
  P(1, 0) += q.s; P(2, 1) += q.s; P(3, 2) += q.s;
  for (double &v : P.d) v *= 0.5;
  return P;
}
// left_mult_mat (quaternion.cpp:339-358)
Mat left_mult_mat(const Quat &q) {
  Mat L(4, 4);
  L(0, 0) = q.s;
  for (int k = 0; k < 3; k++) { L(0, 1 + k) = -q.v[k]; L(1 + k, 0) = q.v[k]; }
  const M3 rc = rcross(q.v);
  for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) L(1 + i, 1 + j) = -rc(i, j);
  L(1, 1) += q.s; L(2, 2) += q.s; L(3, 3) += q.s;
  return L;
}

// filament.cpp:8-55
M3 bfm_lie_derivative(const Quat &q1, const Quat &q2, bool upper, double DL, double KB,
                      double KT) {
  Mat A = upper ? psi_mat(q2) : psi_mat(q1);
  const double s1 = upper ? 1.0 / DL : -1.0 / DL;
  for (double &v : A.d) v *= s1;
  Mat dMdu = matmul(left_mult_mat(qconj(midpoint(q1, q2))), A);
  const double s2 = upper ? 0.5 * DL : -0.5 * DL;
  for (double &v : A.d) v *= s2;
  const Mat t = matmul(left_mult_mat(qscale(qsub(q2, q1), 1.0 / DL)), A);
  for (size_t k = 0; k < dMdu.d.size(); k++) dMdu.d[k] += t.d[k];
  M3 out;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) out(i, j) = 2.0 * dMdu(1 + i, j) * (i == 0 ? KT : KB);
  return out;
}

// filament.cpp:1112-1220 (instability branch: end force at the tip, the
// perturbations at nt = 0 when `perturb`)
void internal_forces(Filament &F, double DL, double KB, double KT, double end_force,
                     bool perturb) {
  const int S = int(F.seg.size());
  double *f = F.f;
  std::fill(f, f + 6 * S, 0.0);
  double t[3], tc[3];
  for (int i = 0; i < S - 1; i++) {
    const double *lam = &F.lam[3 * i];
    for (int k = 0; k < 3; k++) { f[6 * i + k] -= lam[k]; f[6 * (i + 1) + k] += lam[k]; }
    tangent(F.seg[i].q, t);
    cross(t, lam, tc);
    for (int k = 0; k < 3; k++) f[6 * i + 3 + k] -= 0.5 * DL * tc[k];
    tangent(F.seg[i + 1].q, t);
    cross(t, lam, tc);
    for (int k = 0; k < 3; k++) f[6 * (i + 1) + 3 + k] -= 0.5 * DL * tc[k];
    const M3 R = rot_mat(midpoint(F.seg[i].q, F.seg[i + 1].q));
    double Mb[3], M[3];
    body_frame_moment(F, i, DL, KB, KT, Mb);
    mv3(R, Mb, M);
    for (int k = 0; k < 3; k++) { f[6 * i + 3 + k] += M[k]; f[6 * (i + 1) + 3 + k] -= M[k]; }
  }
  for (int k = 0; k < 3; k++) f[k] += F.teth[k];
  if (end_force != 0.0) {
    tangent(F.seg[S - 1].q, t);
    for (int k = 0; k < 3; k++) f[6 * (S - 1) + k] -= end_force * t[k];
    if (perturb)
      for (int k = 0; k < 3; k++) {
        f[6 * 4 + k] += F.pert[k];
        f[6 * 12 + k] += F.pert[3 + k];
      }
  }
  double ct[3];
  dexpinv_transpose(ct, F.seg[0].u, F.clamp);
  for (int k = 0; k < 3; k++) f[3 + k] += ct[k];
}

M3 elastic_rc(const double *K, const Quat &q) {
  double t[3], n[3], b[3];
  tangent(q, t); normal(q, n); binormal(q, b);
  return K[0] * rcross(t) + (K[1] * rcross(n) + K[2] * rcross(b));
}

// filament.cpp:1222-1404
Mat jacobian_lie_block(const Filament &F, int nt, const orc_sim_params &p, double DL) {
  const int S = int(F.seg.size());
  Mat E(3 * S, 3 * S);
  const double tsf = (nt < p.num_euler_steps) ? p.dt : 2.0 * p.dt / 3.0;
  const double Tfac = 1.0 / (8.0 * kPi * p.mu * p.a_seg * p.a_seg * p.a_seg);
  auto diag_block = [&](const M3 &dOmega, int k) {
    const M3 Lie = -1.0 * rcross(F.seg[k].u);
    const double aO[3] = {Tfac * F.f[6 * k + 3], Tfac * F.f[6 * k + 4], Tfac * F.f[6 * k + 5]};
    const M3 Om = rcross(aO);
    double Ou[3];
    mv3(Om, F.seg[k].u, Ou);
    return -tsf * (dOmega - 0.5 * (Om + Lie * dOmega) +
                   (1.0 / 12.0) * (rcross(Ou) + Lie * Om + Lie * Lie * dOmega));
  };
  auto off_block = [&](const M3 &dOmega, int k) {
    const M3 Lie = -1.0 * rcross(F.seg[k].u);
    return -tsf * (dOmega - 0.5 * (Lie * dOmega) + (1.0 / 12.0) * (Lie * Lie * dOmega));
  };
  double t[3];
  // first segment
  {
    tangent(F.seg[0].q, t);
    const M3 Cp = -0.5 * DL * (rcross(&F.lam[0]) * rcross(t));
    double Kph[3];
    body_frame_moment(F, 0, DL, p.kb, p.kt, Kph);
    const M3 R = rot_mat(midpoint(F.seg[0].q, F.seg[1].q));
    M3 El = 0.5 * elastic_rc(Kph, F.seg[0].q) +
            R * bfm_lie_derivative(F.seg[0].q, F.seg[1].q, false, DL, p.kb, p.kt);
    E.set3(0, 0, diag_block(Tfac * (El + Cp), 0));
    El = 0.5 * elastic_rc(Kph, F.seg[1].q) +
         R * bfm_lie_derivative(F.seg[0].q, F.seg[1].q, true, DL, p.kb, p.kt);
    E.set3(0, 3, off_block(Tfac * El, 0));
  }
  // last segment
  {
    const int e = S - 1;
    tangent(F.seg[e].q, t);
    const M3 Cp = -0.5 * DL * (rcross(&F.lam[3 * (S - 2)]) * rcross(t));
    double Kmh[3];
    body_frame_moment(F, S - 2, DL, p.kb, p.kt, Kmh);
    const M3 R = rot_mat(midpoint(F.seg[S - 2].q, F.seg[e].q));
    M3 El = -0.5 * elastic_rc(Kmh, F.seg[e].q) -
            R * bfm_lie_derivative(F.seg[S - 2].q, F.seg[e].q, true, DL, p.kb, p.kt);
    E.set3(3 * e, 3 * e, diag_block(Tfac * (El + Cp), e));
    El = -0.5 * elastic_rc(Kmh, F.seg[S - 2].q) -
         R * bfm_lie_derivative(F.seg[S - 2].q, F.seg[e].q, false, DL, p.kb, p.kt);
    E.set3(3 * e, 3 * e - 3, off_block(Tfac * El, e));
  }
  for (int k = 1; k < S - 1; k++) {
    tangent(F.seg[k].q, t);
    double Kph[3], Kmh[3];
    body_frame_moment(F, k, DL, p.kb, p.kt, Kph);
    body_frame_moment(F, k - 1, DL, p.kb, p.kt, Kmh);
    const Quat qph = midpoint(F.seg[k].q, F.seg[k + 1].q);
    const Quat qmh = midpoint(F.seg[k - 1].q, F.seg[k].q);
    const double lv[3] = {F.lam[3 * (k - 1)] + F.lam[3 * k], F.lam[3 * (k - 1) + 1] + F.lam[3 * k + 1],
                          F.lam[3 * (k - 1) + 2] + F.lam[3 * k + 2]};
    const M3 Cp = -0.5 * DL * (rcross(lv) * rcross(t));
    const M3 Rp = rot_mat(qph), Rm = rot_mat(qmh);
    M3 El = 0.5 * elastic_rc(Kph, F.seg[k].q) +
            Rp * bfm_lie_derivative(F.seg[k].q, F.seg[k + 1].q, false, DL, p.kb, p.kt);
    El = El - 0.5 * elastic_rc(Kmh, F.seg[k].q);
    El = El - Rm * bfm_lie_derivative(F.seg[k - 1].q, F.seg[k].q, true, DL, p.kb, p.kt);
    E.set3(3 * k, 3 * k, diag_block(Tfac * (Cp + El), k));
    El = -0.5 * elastic_rc(Kmh, F.seg[k - 1].q) -
         Rm * bfm_lie_derivative(F.seg[k - 1].q, F.seg[k].q, false, DL, p.kb, p.kt);
    E.set3(3 * k, 3 * k - 3, off_block(Tfac * El, k));
    El = 0.5 * elastic_rc(Kph, F.seg[k + 1].q) +
         Rp * bfm_lie_derivative(F.seg[k].q, F.seg[k + 1].q, true, DL, p.kb, p.kt);
    E.set3(3 * k, 3 * k + 3, off_block(Tfac * El, k));
  }
  for (int k = 0; k < 3 * S; k++) E(k, k) += 1.0;
  return E;
}

// filament.cpp:1406-1539; returns false when singular
bool invert_approx_jacobian(Filament &F, int nt, const orc_sim_params &p, double DL) {
  const int S = int(F.seg.size());
  Mat J(6 * S, 6 * S);
  const double mu = p.mu, a = p.a_seg;
  double fac = (nt < p.num_euler_steps) ? -p.dt / (6.0 * kPi * mu * a) : -p.dt / (9.0 * kPi * mu * a);
  for (int i = 0; i < 3 * S; i++) {
    J(i, i) = fac;
    if (i < 3 * S - 3) J(i, i + 3) = -fac;
  }
  double t[3];
  for (int j = 1; j < S; j++) {
    const int jpos = 3 * (S + j);
    tangent(F.seg[j].q, t);
    const M3 tm = DL * rcross(t);
    J.set3(3 * j, jpos, 0.5 * tm);
    for (int i = j + 1; i < S; i++) J.set3(3 * i, jpos, tm);
  }
  if (p.end_force != 0.0) {  // filament.cpp:1466-1470
    tangent(F.seg[S - 1].q, t);
    J.add3(3 * (S - 1), 3 * (2 * S - 1), -(fac * p.end_force) * rcross(t));
  }
  fac = (nt < p.num_euler_steps) ? DL * p.dt / (16.0 * kPi * mu * a * a * a)
                                 : DL * p.dt / (24.0 * kPi * mu * a * a * a);
  for (int j = 1; j < S; j++) {
    const int jpos = 3 * j;
    int ipos = 3 * (S + j - 1);
    tangent(F.seg[j - 1].q, t);
    M3 um = rcross(F.seg[j - 1].u);
    M3 tm = -1.0 * rcross(t);
    J.set3(ipos, jpos, fac * (tm + 0.5 * (um * tm) + (1.0 / 12.0) * (um * um * tm)));
    ipos += 3;
    tangent(F.seg[j].q, t);
    um = rcross(F.seg[j].u);
    tm = -1.0 * rcross(t);
    J.set3(ipos, jpos, fac * (tm + 0.5 * (um * tm) + (1.0 / 12.0) * (um * um * tm)));
  }
  const Mat E = jacobian_lie_block(F, nt, p, DL);
  for (int i = 0; i < 3 * S; i++)
    for (int j = 0; j < 3 * S; j++) J(3 * S + i, 3 * S + j) = E(i, j);
  fac = (nt < p.num_euler_steps) ? -p.dt / (8.0 * kPi * mu * a * a * a)
                                 : -p.dt / (12.0 * kPi * mu * a * a * a);
  const M3 um = rcross(F.seg[0].u);
  const M3 blk = eye3() + 0.5 * um + (1.0 / 12.0) * (um * um);
  const M3 blkT = eye3() - 0.5 * um + (1.0 / 12.0) * (um * um);
  F.ecb1 = J.get3(3 * S, 3 * S);
  F.ecb2 = J.get3(3 * S + 3, 3 * S);
  J.set3(3 * S, 3 * S, fac * (blk * blkT));
  J.set3(3 * S + 3, 3 * S, M3());
  if (!invert(J)) return false;
  F.Jinv = J;
  return true;
}

// filament.cpp:1541-1575
void fil_update(Filament &F, const double *u, double DL) {
  const int S = int(F.seg.size());
  for (int k = 0; k < 3; k++) F.teth[k] += u[k];
  for (int i = 1; i < S; i++) {
    for (int k = 0; k < 3; k++) F.lam[3 * (i - 1) + k] += u[3 * i + k];
    seg_update(F.seg[i], &u[3 * (S + i)]);
  }
  for (int k = 0; k < 3; k++) F.clamp[k] += u[3 * S + k];
  robot_arm(F, DL);
}

// ------------------------------------------------------------------ swimmer

void body_blob_positions(const Swimmer &B, double *xb) {
  const M3 R = rot_mat(B.q);
  for (int m = 0; m < B.nblob; m++) {
    double r[3];
    mv3(R, &B.blob_ref[3 * m], r);
    for (int k = 0; k < 3; k++) xb[3 * m + k] = B.x[k] + r[k];
  }
}

void base_pos(const Swimmer &B, const M3 &R, int i, double *out) {
  double r[3];
  mv3(R, &B.fil_ref[3 * i], r);
  for (int k = 0; k < 3; k++) out[k] = B.x[k] + r[k];
}

// rigid_body.cpp:247-332
void body_initial_guess(Swimmer &B, int nt, int neuler) {
  double gx[3], gu[3];
  for (int k = 0; k < 3; k++) {
    if (nt <= neuler) { gx[k] = 2.0 * B.x[k] - B.xm1[k]; gu[k] = B.u[k]; }
    else { gx[k] = 3.0 * (B.x[k] - B.xm1[k]) + B.xm2[k]; gu[k] = 2.0 * B.u[k] - B.um1[k]; }
  }
  std::memcpy(B.xm2, B.xm1, sizeof B.xm2);
  std::memcpy(B.xm1, B.x, sizeof B.xm1);
  std::memcpy(B.x, gx, sizeof gx);
  std::memcpy(B.um1, B.u, sizeof B.um1);
  std::memcpy(B.u, gu, sizeof gu);
  B.qm1 = B.q;
  B.q = qmul(lie_exp(B.u), B.qm1);
  B.Qinit = rot_mat(B.q);
}

// swimmer.cpp:240-274
void sw_initial_guess(Swimmer &B, int nt, int neuler, double DL) {
  body_initial_guess(B, nt, neuler);
  const M3 R = rot_mat(B.q);
  for (int i = 0; i < B.nfil; i++) {
    double bp[3];
    base_pos(B, R, i, bp);
    fil_initial_guess(B.fil[i], nt, bp, B.u, neuler, DL);
  }
}

// swimmer.cpp:276-325
void sw_forces(Swimmer &B, const orc_sim_params &p, double DL, int nt) {
  std::fill(B.f, B.f + 6, 0.0);
  for (Filament &F : B.fil) {
    internal_forces(F, DL, p.kb, p.kt, p.end_force, nt == 0 && p.perturbation);
    const int S = int(F.seg.size());
    for (int m = 0; m < S; m++) {
      const double *fs = &F.f[6 * m];
      for (int k = 0; k < 3; k++) B.f[k] -= fs[k];
      const double d[3] = {F.seg[m].x[0] - B.x[0], F.seg[m].x[1] - B.x[1], F.seg[m].x[2] - B.x[2]};
      double c[3];
      cross(d, fs, c);
      for (int k = 0; k < 3; k++) B.f[3 + k] -= fs[3 + k] + c[k];
    }
  }
}

Mat rotate66(const Mat &N, const M3 &Q) {
  Mat out(6, 6);
  const M3 QT = transpose(Q);
  for (int bi = 0; bi < 2; bi++)
    for (int bj = 0; bj < 2; bj++) out.set3(3 * bi, 3 * bj, Q * N.get3(3 * bi, 3 * bj) * QT);
  return out;
}

// A matrix (6S x 6) of swimmer.cpp:535-541,673-680 / 806-831
Mat fil_A(const Swimmer &B, int i, const Filament &F, double DL) {
  const int S = int(F.seg.size());
  Mat A(6 * S, 6);
  for (int n = 0; n < S; n++) { A(3 * n, 0) = 1; A(3 * n + 1, 1) = 1; A(3 * n + 2, 2) = 1; }
  A.set3(3 * S, 3, F.ecb1);
  A.set3(3 * S + 3, 3, F.ecb2);
  double d[3];
  mv3(B.Qinit, &B.fil_ref[3 * i], d);
  const double dn = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  double v[3];
  for (int k = 0; k < 3; k++) v[k] = (0.5 * DL + dn) * d[k] / dn;
  const M3 Dm = rcross(v);
  for (int n = 0; n < S; n++) A.set3(3 * n, 3, Dm);
  return A;
}

// swimmer.cpp:329-688 (GMRES branch 466-682, no end force)
bool sw_prepare_jacobian_inv(Swimmer &B, int nt, const orc_sim_params &p, double DL) {
  for (Filament &F : B.fil)
    if (!invert_approx_jacobian(F, nt, p, DL)) return false;
  const int S = int(B.fil[0].seg.size());
  const double dtf = (nt < p.num_euler_steps) ? -p.dt : -2.0 * p.dt / 3.0;
  const M3 blc = -1.0 * rcross(B.u);
  const M3 cbo = (-1.0 / dtf) * blc;
  Mat schur(6, 6);
  for (int k = 0; k < 6; k++) schur(k, k) = 1.0;
  schur.add3(3, 3, dtf * ((1.0 / 12.0) * (blc * cbo) - 0.5 * cbo));
  M3 old;  // omega_lie_deriv
  for (int i = 0; i < B.nfil; i++) {
    const Filament &F = B.fil[i];
    double d[3];
    mv3(B.Qinit, &B.fil_ref[3 * i], d);
    old = old - rcross(F.teth) * rcross(d);
    double ucl[3];
    cross(B.u, F.clamp, ucl);
    old = old + 0.5 * rcross(F.clamp) - (1.0 / 12.0) * (rcross(ucl) + blc * rcross(F.clamp));
    if (p.end_force != 0.0) {  // swimmer.cpp:497-503
      double t[3], v[3];
      tangent(F.seg[S - 1].q, t);
      const double dn = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      for (int k = 0; k < 3; k++) { t[k] *= p.end_force; v[k] = (0.5 * DL + dn) * d[k] / dn; }
      old = old + rcross(t) * rcross(v);
    }
  }
  const Mat bm = rotate66(B.Nref, B.Qinit);
  old = bm.get3(3, 3) * old;
  schur.add3(3, 3, dtf * (old - 0.5 * (blc * old) + (1.0 / 12.0) * (blc * blc * old)));
  const M3 Dinv = eye3() + (1.0 / 12.0) * (blc * blc) - 0.5 * blc;
  const M3 DinvT = eye3() + (1.0 / 12.0) * (blc * blc) + 0.5 * blc;
  const double R = B.R;
  const double mu = p.mu, as = p.a_seg;
  B.Bblk.assign(B.nfil, Mat());
  for (int i = 0; i < B.nfil; i++) {
    const Filament &F = B.fil[i];
    Mat Bm(6, 6 * S);
    Bm(0, 0) = -1; Bm(1, 1) = -1; Bm(2, 2) = -1;
    double d[3];
    mv3(B.Qinit, &B.fil_ref[3 * i], d);
    Bm.set3(3, 0, rcross(d));
    Bm.set3(3, 3 * S, -1.0 * DinvT);
    if (p.end_force != 0.0) {  // swimmer.cpp:557-571
      double t[3], ft[3], arm[3];
      tangent(F.seg[S - 1].q, t);
      for (int k = 0; k < 3; k++) {
        ft[k] = p.end_force * t[k];
        arm[k] = 0.5 * DL * t[k] - (F.seg[S - 1].x[k] - B.x[k]);
      }
      const M3 Ef = rcross(ft);
      Bm.set3(0, 6 * S - 3, Ef);
      Bm.set3(3, 6 * S - 3, rcross(arm) * Ef);
      for (int m = 1; m < S - 1; m++) {
        double tmv[3];
        tangent(F.seg[m].q, tmv);
        for (int k = 0; k < 3; k++) tmv[k] *= DL;
        Bm.set3(3, 3 * (S + m), Ef * rcross(tmv));
      }
    }
    Bm = matmul(bm, Bm);
    for (int n = 0; n < S; n++) {
      double rh[3] = {F.seg[n].x[0] - B.x[0], F.seg[n].x[1] - B.x[1], F.seg[n].x[2] - B.x[2]};
      const double r = std::sqrt(rh[0] * rh[0] + rh[1] * rh[1] + rh[2] * rh[2]);
      for (int k = 0; k < 3; k++) rh[k] /= r;
      M3 rr;
      for (int a = 0; a < 3; a++) for (int b = 0; b < 3; b++) rr(a, b) = rh[a] * rh[b];
      const M3 I = eye3();
      const M3 Mtt = (1.0 / (8.0 * kPi * mu * r)) *
                     ((1.0 + (R * R + as * as) / (3.0 * r * r)) * I + (1.0 - (R * R + as * as) / (r * r)) * rr);
      const M3 Mtr = (1.0 / (8.0 * kPi * mu * r * r)) * rcross(rh);
      const M3 Mrr = (1.0 / (16.0 * kPi * mu * r * r * r)) * (3.0 * rr - I);
      double t[3];
      tangent(F.seg[n].q, t);
      const M3 tm = rcross(t);
      if (n == 0) {
        Bm.add3(0, 0, Mtt);
        Bm.add3(3, 0, Mtr);
        Bm.add3(0, 3 * S, Mtr * DinvT);
        Bm.add3(3, 3 * S, Mrr * DinvT);
        Bm.add3(0, 3, 0.5 * DL * (Mtr * tm) - Mtt);
        Bm.add3(3, 3, 0.5 * DL * (Mrr * tm) - Mtr);
      } else if (n == S - 1) {
        if (p.end_force != 0.0) {  // swimmer.cpp:630-633
          Bm.add3(0, 6 * S - 3, -p.end_force * (Mtt * tm));
          Bm.add3(3, 6 * S - 3, -p.end_force * (Mtr * tm));
        }
        Bm.add3(0, 3 * S - 3, 0.5 * DL * (Mtr * tm) + Mtt);
        Bm.add3(3, 3 * S - 3, 0.5 * DL * (Mrr * tm) + Mtr);
      } else {
        Bm.add3(0, 3 * n, 0.5 * DL * (Mtr * tm) + Mtt);
        Bm.add3(3, 3 * n, 0.5 * DL * (Mrr * tm) + Mtr);
        Bm.add3(0, 3 * (n + 1), 0.5 * DL * (Mtr * tm) - Mtt);
        Bm.add3(3, 3 * (n + 1), 0.5 * DL * (Mrr * tm) - Mtr);
      }
    }
    for (int n = 0; n < S; n++) {
      Bm.set3(3, 3 * n, Dinv * Bm.get3(3, 3 * n));
      Bm.set3(3, 3 * (n + S), Dinv * Bm.get3(3, 3 * (n + S)));
    }
    for (double &v : Bm.d) v *= dtf;
    B.Bblk[i] = matmul(Bm, F.Jinv);
    const Mat A = fil_A(B, i, F, DL);
    const Mat BA = matmul(B.Bblk[i], A);
    for (int k = 0; k < 36; k++) schur.d[k] -= BA.d[k];
  }
  if (!invert(schur)) return false;
  B.schur_inv = schur;
  return true;
}

// swimmer.cpp:793-839
void sw_jacobian_inv_mult(const Swimmer &B, const double *in, double *out, double DL) {
  const int S = int(B.fil[0].seg.size());
  const int nf = 6 * S;
  double b2[6];
  for (int k = 0; k < 6; k++) b2[k] = in[B.nfil * nf + k];
  for (int i = 0; i < B.nfil; i++) {
    const Mat &Bb = B.Bblk[i];
    for (int r = 0; r < 6; r++) {
      double s = 0;
      for (int c = 0; c < nf; c++) s += Bb(r, c) * in[i * nf + c];
      b2[r] -= s;
    }
  }
  double ob[6];
  for (int r = 0; r < 6; r++) {
    double s = 0;
    for (int c = 0; c < 6; c++) s += B.schur_inv(r, c) * b2[c];
    ob[r] = s;
  }
  for (int k = 0; k < 6; k++) out[B.nfil * nf + k] = ob[k];
  std::vector<double> tmp(nf);
  for (int i = 0; i < B.nfil; i++) {
    const Mat A = fil_A(B, i, B.fil[i], DL);
    for (int r = 0; r < nf; r++) {
      double s = 0;
      for (int c = 0; c < 6; c++) s += A(r, c) * ob[c];
      tmp[r] = in[i * nf + r] - s;
    }
    const Mat &Ji = B.fil[i].Jinv;
    for (int r = 0; r < nf; r++) {
      double s = 0;
      for (int c = 0; c < nf; c++) s += Ji(r, c) * tmp[c];
      out[i * nf + r] = s;
    }
  }
}

// swimmer.cpp:847-871 + rigid_body.cpp:354-381
void sw_update(Swimmer &B, const double *up, double DL) {
  const int S = int(B.fil[0].seg.size());
  const double *bu = &up[6 * B.nfil * S];
  for (int k = 0; k < 3; k++) { B.x[k] += bu[k]; B.u[k] += bu[3 + k]; }
  B.q = qmul(lie_exp(B.u), B.qm1);
  const M3 R = rot_mat(B.q);
  for (int i = 0; i < B.nfil; i++) {
    double bp[3];
    base_pos(B, R, i, bp);
    accept_state_from_rigid_body(B.fil[i], bp, B.u, DL);
    fil_update(B.fil[i], &up[6 * S * i], DL);
  }
}

// ------------------------------------------------------- mobility solver

// The pair sums over one group of particles: segments [seg0, seg0 + ns),
// blobs [blob0, blob0 + nb).
void pair_apply_group(orc_sim *s, int seg0, int ns, int blob0, int nb, const double *Fs,
                      const double *Fb, bool ss, bool sb, bool bs, bool bb, double *Vs,
                      double *Vb) {
  const double *xs = &s->xs[3 * seg0], *xb = &s->xb[3 * blob0];
  double *vs = Vs ? &Vs[6 * seg0] : nullptr;
  double *vb = Vb ? &Vb[3 * blob0] : nullptr;
  if (ss) orc_Mss(vs, &Fs[6 * seg0], xs, ns, 0, ns, s->p.a_seg, s->p.mu, s->nthreads);
  if (sb) orc_Msb(vs, &Fb[3 * blob0], xs, ns, xb, nb, 0, ns, s->p.a_seg, s->p.a_blob, s->p.mu,
                  s->nthreads);
  if (bs) orc_Mbs(vb, &Fs[6 * seg0], xb, nb, xs, ns, 0, nb, s->p.a_seg, s->p.a_blob, s->p.mu,
                  s->nthreads);
  if (bb) {
    if (bs) {
      std::vector<double> t(3 * nb);
      orc_Mbb(t.data(), &Fb[3 * blob0], xb, nb, 0, nb, s->p.a_blob, s->p.mu, s->nthreads);
      for (int k = 0; k < 3 * nb; k++) vb[k] += t[k];
    } else {
      orc_Mbb(vb, &Fb[3 * blob0], xb, nb, 0, nb, s->p.a_blob, s->p.mu, s->nthreads);
    }
  }
}

// cuda_functions.cu:246-863 over all particles (coupled swimmers) or per
// swimmer (independent swimmers)
void pair_apply_segs_blobs(orc_sim *s, const double *Fs, const double *Fb, bool ss, bool sb,
                           bool bs, bool bb, double *Vs, double *Vb) {
  if (s->p.interact) {
    pair_apply_group(s, 0, s->Ns, 0, s->Nb, Fs, Fb, ss, sb, bs, bb, Vs, Vb);
    return;
  }
  for (const Swimmer &B : s->sw)
    pair_apply_group(s, B.seg0, B.nfil * s->S, B.blob0, B.nblob, Fs, Fb, ss, sb, bs, bb, Vs, Vb);
}

// mobility_solver.cpp:132-192
void read_positions_and_forces(orc_sim *s) {
  for (const Swimmer &B : s->sw) body_blob_positions(B, &s->xb[3 * B.blob0]);
  s->fs_dev = s->fs;
  std::fill(s->frep.begin(), s->frep.end(), 0.0);
  if (s->p.interact) {
    orc_barrier(s->fs_dev.data(), s->frep.data(), s->xs.data(), s->Ns, s->xb.data(), s->Nb, s->S,
                s->p.a_seg, s->p.a_blob, s->p.barrier_K, s->nthreads);
  } else {
    for (const Swimmer &B : s->sw) {
      const int ns = B.nfil * s->S;
      orc_barrier(&s->fs_dev[6 * B.seg0], &s->frep[3 * B.blob0], &s->xs[3 * B.seg0], ns,
                  &s->xb[3 * B.blob0], B.nblob, s->S, s->p.a_seg, s->p.a_blob, s->p.barrier_K,
                  s->nthreads);
    }
  }
  for (Swimmer &B : s->sw) {
    const M3 R = rot_mat(B.q);
    for (int m = 0; m < B.nblob; m++) {
      const double *fr = &s->frep[3 * (B.blob0 + m)];
      double r[3], c[3];
      mv3(R, &B.blob_ref[3 * m], r);
      cross(r, fr, c);
      for (int k = 0; k < 3; k++) { B.f[k] += fr[k]; B.f[3 + k] += c[k]; }
    }
  }
}

// index of swimmer n's body velocity in the GMRES vector
inline int body_slot(const orc_sim *s, int n) { return 3 * s->Nb + 6 * n; }

// mobility_solver.cpp:651-727 (blob part scaled once per blob)
void apply_preconditioner(orc_sim *s, const std::vector<double> &in, std::vector<double> &out) {
  const double Minv = 6.0 * kPi * s->p.mu * s->p.a_blob;
  out.assign(in.size(), 0.0);
  for (size_t n = 0; n < s->sw.size(); n++) {
    const Swimmer &B = s->sw[n];
    const M3 Q = rot_mat(B.q);
    const Mat bm = rotate66(B.Nref, Q), kk = rotate66(B.KTKinv, Q);
    double b[6] = {0, 0, 0, 0, 0, 0};
    for (int m = 0; m < B.nblob; m++) {
      const double *v = &in[3 * (B.blob0 + m)];
      double d[3], c[3];
      mv3(Q, &B.blob_ref[3 * m], d);
      cross(d, v, c);
      for (int k = 0; k < 3; k++) { b[k] -= v[k]; b[3 + k] -= c[k]; }
    }
    const int vid = body_slot(s, int(n));
    double ou[6];
    for (int r = 0; r < 6; r++) {
      double a = 0, c = 0;
      for (int k = 0; k < 6; k++) { a += kk(r, k) * b[k]; c += bm(r, k) * in[vid + k]; }
      ou[r] = a - c;
    }
    for (int k = 0; k < 6; k++) out[vid + k] = ou[k];
    for (int m = 0; m < B.nblob; m++) {
      const int id = 3 * (B.blob0 + m);
      double d[3], w[3];
      mv3(Q, &B.blob_ref[3 * m], d);
      cross(&ou[3], d, w);
      for (int k = 0; k < 3; k++) out[id + k] = Minv * (in[id + k] + ou[k] + w[k]);
    }
  }
}

// mobility_solver.cpp:905-964
void system_matrix_mult(orc_sim *s, const std::vector<double> &in, std::vector<double> &out) {
  out.assign(in.size(), 0.0);
  std::copy(in.begin(), in.begin() + 3 * s->Nb, s->fb.begin());
  std::vector<double> v(3 * s->Nb, 0.0);
  pair_apply_segs_blobs(s, nullptr, s->fb.data(), false, false, false, true, nullptr, v.data());
  for (size_t n = 0; n < s->sw.size(); n++) {
    const Swimmer &B = s->sw[n];
    const M3 Q = rot_mat(B.q);
    const int sid = body_slot(s, int(n));
    for (int m = 0; m < B.nblob; m++) {
      const int id = 3 * (B.blob0 + m);
      double d[3], c[3], w[3];
      mv3(Q, &B.blob_ref[3 * m], d);
      cross(d, &in[id], c);
      for (int k = 0; k < 3; k++) { out[sid + k] -= in[id + k]; out[sid + 3 + k] -= c[k]; }
      cross(&in[sid + 3], d, w);
      for (int k = 0; k < 3; k++) out[id + k] -= in[sid + k] + w[k];
    }
  }
  for (int k = 0; k < 3 * s->Nb; k++) out[k] += v[k];
}

double vnorm(const std::vector<double> &a) {
  double s = 0;
  for (double v : a) s += v * v;
  return std::sqrt(s);
}

// mobility_solver.cpp:970-1354 (right preconditioning)
int solve_linear_system(orc_sim *s) {
  const int n = s->nsys;
  const double tol = s->p.gmres_tol;
  const double nrhs = vnorm(s->rhs);
  if (nrhs == 0.0) {
    std::fill(s->fb.begin(), s->fb.end(), 0.0);
    std::fill(s->v_bodies.begin(), s->v_bodies.end(), 0.0);
    return 0;
  }
  std::vector<double> soln(n);
  std::copy(s->fb.begin(), s->fb.end(), soln.begin());
  std::copy(s->v_bodies.begin(), s->v_bodies.end(), soln.begin() + 3 * s->Nb);
  const int maxit = std::min(n, s->p.gmres_max_iter);
  std::vector<std::vector<double>> Q(maxit + 1, std::vector<double>(n, 0.0));
  std::vector<double> H(size_t(maxit + 1) * (maxit + 1), 0.0), beta(maxit + 1, 0.0), SN(maxit),
      CS(maxit);
  auto h = [&](int i, int j) -> double & { return H[size_t(i) * (maxit + 1) + j]; };
  std::vector<double> tmp, tmp2;
  system_matrix_mult(s, soln, tmp);
  for (int k = 0; k < n; k++) Q[0][k] = s->rhs[k] - tmp[k];
  beta[0] = vnorm(Q[0]);
  if (std::fabs(beta[0]) <= tol * nrhs) return 0;
  if (std::fabs(beta[0]) > nrhs) {
    std::fill(soln.begin(), soln.end(), 0.0);
    beta[0] = nrhs;
    Q[0] = s->rhs;
  }
  for (double &v : Q[0]) v /= beta[0];
  for (int it = 1; it <= maxit; it++) {
    apply_preconditioner(s, Q[it - 1], tmp);
    system_matrix_mult(s, tmp, Q[it]);
    for (int i = 0; i < it; i++) {
      double d = 0;
      for (int k = 0; k < n; k++) d += Q[it][k] * Q[i][k];
      h(i, it - 1) = d;
      for (int k = 0; k < n; k++) Q[it][k] -= d * Q[i][k];
    }
    const double qn = vnorm(Q[it]);
    h(it, it - 1) = qn;
    for (double &v : Q[it]) v /= qn;
    for (int i = 1; i < it; i++) {
      const double t = CS[i - 1] * h(i - 1, it - 1) + SN[i - 1] * h(i, it - 1);
      h(i, it - 1) = CS[i - 1] * h(i, it - 1) - SN[i - 1] * h(i - 1, it - 1);
      h(i - 1, it - 1) = t;
    }
    if (h(it, it - 1) == 0.0) {
      CS[it - 1] = 1.0; SN[it - 1] = 0.0;
    } else if (std::fabs(h(it, it - 1)) > std::fabs(h(it - 1, it - 1))) {
      const double t = h(it - 1, it - 1) / h(it, it - 1);
      SN[it - 1] = (2.0 * double(h(it, it - 1) > 0) - 1.0) / std::sqrt(1.0 + t * t);
      CS[it - 1] = t * SN[it - 1];
    } else {
      const double t = h(it, it - 1) / h(it - 1, it - 1);
      CS[it - 1] = (2.0 * double(h(it - 1, it - 1) > 0) - 1.0) / std::sqrt(1.0 + t * t);
      SN[it - 1] = t * CS[it - 1];
    }
    h(it - 1, it - 1) = CS[it - 1] * h(it - 1, it - 1) + SN[it - 1] * h(it, it - 1);
    h(it, it - 1) = 0.0;
    beta[it] = -SN[it - 1] * beta[it - 1];
    beta[it - 1] = CS[it - 1] * beta[it - 1];
    const double rel = std::fabs(beta[it]) / nrhs;
    if (rel <= tol || it == maxit) {
      std::vector<double> y(it, 0.0);
      y[it - 1] = beta[it - 1] / h(it - 1, it - 1);
      for (int i = it - 2; i >= 0; i--) {
        y[i] = beta[i];
        for (int j = it - 1; j > i; j--) y[i] -= h(i, j) * y[j];
        y[i] /= h(i, i);
      }
      std::vector<double> ts(n, 0.0);
      for (int i = 0; i < it; i++)
        for (int k = 0; k < n; k++) ts[k] += y[i] * Q[i][k];
      apply_preconditioner(s, ts, tmp2);
      for (int k = 0; k < n; k++) soln[k] += tmp2[k];
      std::copy(soln.begin(), soln.begin() + 3 * s->Nb, s->fb.begin());
      std::copy(soln.begin() + 3 * s->Nb, soln.end(), s->v_bodies.begin());
      return it;
    }
  }
  return 0;
}

// mobility_solver.cpp:1358-1402
int compute_velocities(orc_sim *s) {
  read_positions_and_forces(s);
  std::fill(s->vs.begin(), s->vs.end(), 0.0);
  pair_apply_segs_blobs(s, s->fs_dev.data(), nullptr, true, false, true, false, s->vs.data(),
                        s->vb.data());
  // assemble_rhs (mobility_solver.cpp:361-421)
  for (int k = 0; k < 3 * s->Nb; k++) s->rhs[k] = -s->vb[k];
  for (size_t n = 0; n < s->sw.size(); n++)
    for (int k = 0; k < 6; k++) s->rhs[body_slot(s, int(n)) + k] = -s->sw[n].f[k];
  const int it = solve_linear_system(s);
  pair_apply_segs_blobs(s, nullptr, s->fb.data(), false, true, false, false, s->vs.data(), nullptr);
  s->last_gmres = it;
  s->total_gmres += it;
  return it;
}

// mobility_solver.cpp:1404-1504,1641-1689 (GMRES branch)
bool compute_errors(orc_sim *s, std::vector<double> &error, int nt) {
  const double c1 = -4.0 / 3.0, c2 = 1.0 / 3.0, c3 = -2.0 / 3.0;
  const double DT = s->p.dt, TOL = s->p.tol, DL = s->DL();
  const int S = s->S;
  bool big = false;
  for (size_t n = 0; n < s->sw.size(); n++) {
    const Swimmer &B = s->sw[n];
    for (int i = 0; i < B.nfil; i++) {
      const Filament &F = B.fil[i];
      for (int j = 0; j < S; j++) {
        const int g = B.seg0 + i * S + j;
        const double *v = &s->vs[6 * g], *om = &s->vs[6 * g + 3];
        int id = B.err0 + 6 * i * S + 3 * j;
        const Segment &sg = F.seg[j];
        for (int k = 0; k < 3; k++) {
          if (nt < s->p.num_euler_steps) error[id + k] = sg.x[k] - sg.xm1[k] - DT * v[k];
          else error[id + k] = sg.x[k] + c1 * sg.xm1[k] + c2 * sg.xm2[k] + c3 * DT * v[k];
        }
        if (!big)
          big = std::fabs(error[id]) > DL * TOL || std::fabs(error[id + 1]) > DL * TOL ||
                std::fabs(error[id + 2]) > DL * TOL;
        id += 3 * S;
        double W[3];
        dexpinv(W, sg.u, om);
        for (int k = 0; k < 3; k++) {
          if (nt < s->p.num_euler_steps) error[id + k] = sg.u[k] - DT * W[k];
          else error[id + k] = sg.u[k] - c2 * sg.um1[k] + c3 * DT * W[k];
        }
        if (!big)
          big = std::fabs(error[id]) > TOL || std::fabs(error[id + 1]) > TOL ||
                std::fabs(error[id + 2]) > TOL;
      }
    }
    const int id = B.err0 + 6 * B.nfil * S;
    const double *U = &s->v_bodies[6 * n];
    double w[3];
    dexpinv(w, B.u, U + 3);
    for (int k = 0; k < 3; k++) {
      if (nt < s->p.num_euler_steps) {
        error[id + k] = B.x[k] - B.xm1[k] - DT * U[k];
        error[id + 3 + k] = B.u[k] - DT * w[k];
      } else {
        error[id + k] = B.x[k] + c1 * B.xm1[k] + c2 * B.xm2[k] + c3 * DT * U[k];
        error[id + 3 + k] = B.u[k] - c2 * B.um1[k] + c3 * DT * w[k];
      }
    }
    if (!big)
      big = std::fabs(error[id]) > TOL * DL || std::fabs(error[id + 1]) > TOL * DL ||
            std::fabs(error[id + 2]) > TOL * DL;
    if (!big)
      big = std::fabs(error[id + 3]) > TOL || std::fabs(error[id + 4]) > TOL ||
            std::fabs(error[id + 5]) > TOL;
  }
  return big;
}

// broyden_solver.cpp:26-68
void find_update(orc_sim *s) {
  for (const Swimmer &B : s->sw) sw_jacobian_inv_mult(B, &s->err[B.err0], &s->upd[B.err0], s->DL());
  const double g = s->p.jacobian_confidence;
  for (double &v : s->upd) v *= -g;
  for (int i = 0; i < s->iter; i++) {
    double d = 0;
    for (int k = 0; k < s->nbroy; k++) d += s->err[k] * s->D(k, i);
    for (int k = 0; k < s->nbroy; k++) s->upd[k] -= d * s->C(k, i);
  }
}

// broyden_solver.cpp:70-130
void end_of_iter(orc_sim *s, bool too_large) {
  if (too_large && s->iter < s->p.max_broyden_iter - 1) {
    for (const Swimmer &B : s->sw)
      sw_jacobian_inv_mult(B, &s->new_err[B.err0], &s->upd[B.err0], s->DL());
    const int it = s->iter;
    const double g = s->p.jacobian_confidence;
    for (int k = 0; k < s->nbroy; k++) s->C(k, it) = -g * s->upd[k];
    for (int i = 0; i < it; i++) {
      double d = 0;
      for (int k = 0; k < s->nbroy; k++) d += s->new_err[k] * s->D(k, i);
      for (int k = 0; k < s->nbroy; k++) s->C(k, it) -= d * s->C(k, i);
    }
    double nn = 0;
    for (int k = 0; k < s->nbroy; k++) {
      s->D(k, it) = s->new_err[k] - s->err[k];
      nn += s->D(k, it) * s->D(k, it);
    }
    nn = std::sqrt(nn);
    for (int k = 0; k < s->nbroy; k++) { s->C(k, it) /= nn; s->D(k, it) /= nn; }
    std::swap(s->err, s->new_err);
  }
  s->iter++;
}

void set_driving(orc_sim *s, int nt) {
  s->nt_cur = nt;
  const double t = (nt + 1) * s->p.dt;
  for (Swimmer &B : s->sw)
    for (Filament &F : B.fil)
      F.st[2] = s->p.curvature_amp * std::sin(2.0 * kPi * t / s->p.curvature_period + F.phase);
}

void all_forces(orc_sim *s) {
  for (Swimmer &B : s->sw) sw_forces(B, s->p, s->DL(), s->nt_cur);
}

}  // namespace

// ----------------------------------------------------------------- C API

extern "C" {

orc_sim *orc_sim_create(const orc_sim_params *p, int nthreads) {
  if (p->perturbation && p->nseg < 13) return nullptr;  // segments 4 and 12
  auto s = std::make_unique<orc_sim>();
  s->p = *p;
  s->S = p->nseg;
  s->nthreads = nthreads > 0 ? nthreads : 1;
  const double DL = 2.2 * p->a_seg;
  int seg0 = 0, blob0 = 0, err0 = 0, fil0 = 0, ref0 = 0;
  s->sw.resize(p->n_swimmers);
  for (int n = 0; n < p->n_swimmers; n++) {
    Swimmer &B = s->sw[n];
    B.nfil = p->nfil[n];
    B.nblob = p->nblob[n];
    B.seg0 = seg0; B.blob0 = blob0; B.err0 = err0;
    B.R = p->sphere_radii ? p->sphere_radii[n] : p->sphere_radius;
    B.blob_ref.assign(p->blob_refs + 3 * ref0, p->blob_refs + 3 * (ref0 + B.nblob));
    B.fil_ref.assign(p->fil_refs + 3 * fil0, p->fil_refs + 3 * (fil0 + B.nfil));
    for (int k = 0; k < 3; k++) B.x[k] = B.xm1[k] = B.xm2[k] = p->body_x ? p->body_x[3 * n + k] : 0.0;
    B.fil.resize(B.nfil);
    B.Nref = Mat(6, 6); B.KTKinv = Mat(6, 6);
    for (int k = 0; k < 6; k++) { B.Nref(k, k) = 1; B.KTKinv(k, k) = 1; }
    seg0 += B.nfil * p->nseg;
    blob0 += B.nblob;
    err0 += 6 * (B.nfil * p->nseg + 1);
    fil0 += B.nfil;
    ref0 += B.nblob;
  }
  s->Ns = seg0; s->Nb = blob0; s->nbroy = err0;
  s->nsys = 3 * s->Nb + 6 * p->n_swimmers;
  s->xs.assign(3 * s->Ns, 0); s->fs.assign(6 * s->Ns, 0); s->vs.assign(6 * s->Ns, 0);
  s->xb.assign(3 * s->Nb, 0); s->fb.assign(3 * s->Nb, 0); s->vb.assign(3 * s->Nb, 0);
  s->frep.assign(3 * s->Nb, 0);
  s->v_bodies.assign(6 * p->n_swimmers, 0);
  s->rhs.assign(s->nsys, 0);
  s->C = Mat(s->nbroy, p->max_broyden_iter);
  s->D = Mat(s->nbroy, p->max_broyden_iter);
  s->err.assign(s->nbroy, 0); s->new_err.assign(s->nbroy, 0); s->upd.assign(s->nbroy, 0);
  // filaments: straight along the base direction (filament.cpp:197-232,297-307)
  int fg = 0;
  for (Swimmer &B : s->sw) {
    for (int i = 0; i < B.nfil; i++, fg++) {
      Filament &F = B.fil[i];
      F.seg.resize(p->nseg);
      F.lam.assign(3 * (p->nseg - 1), 0); F.lamm1 = F.lam; F.lamm2 = F.lam;
      F.f = &s->fs[6 * (B.seg0 + i * p->nseg)];
      F.phase = p->fil_phase ? p->fil_phase[fg] : 0.0;
      if (p->perturbation)
        for (int k = 0; k < 6; k++) F.pert[k] = p->perturbation[6 * fg + k];
      const double *r = &B.fil_ref[3 * i];
      const double rn = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
      const double d[3] = {r[0] / rn, r[1] / rn, r[2] / rn};
      Quat q0;
      q0.s = d[0]; q0.v[0] = 0.0; q0.v[1] = -d[2]; q0.v[2] = d[1];
      q0 = qsqrt(q0);
      for (int j = 0; j < p->nseg; j++) {
        Segment &sg = F.seg[j];
        for (int k = 0; k < 3; k++) sg.x[k] = sg.xm1[k] = sg.xm2[k] = B.x[k] + r[k] + j * DL * d[k];
        sg.q = sg.qm1 = q0;
      }
    }
  }
  return s.release();
}

void orc_sim_destroy(orc_sim *s) { delete s; }

// copies segment positions into the mobility buffer (the reference aliases them)
static void sync_positions(orc_sim *s) {
  for (Swimmer &B : s->sw)
    for (int i = 0; i < B.nfil; i++)
      for (int j = 0; j < s->S; j++)
        for (int k = 0; k < 3; k++)
          s->xs[3 * (B.seg0 + i * s->S + j) + k] = B.fil[i].seg[j].x[k];
}

// mobility_solver.cpp:1693-1771 (per swimmer for independent swimmers)
int orc_sim_make_body_reference_matrices(orc_sim *s) {
  for (Swimmer &B : s->sw) {
    Mat K(6, 6);
    K(0, 0) = K(1, 1) = K(2, 2) = B.nblob;
    for (int m = 0; m < B.nblob; m++) {
      const M3 rc = rcross(&B.blob_ref[3 * m]);
      K.add3(0, 3, rc);
      K.add3(3, 0, -1.0 * rc);
      K.add3(3, 3, -1.0 * (rc * rc));
    }
    if (!invert(K)) return -5;
    B.KTKinv = K;
  }
  sync_positions(s);
  read_positions_and_forces(s);
  for (int n = 0; n < 6; n++) {
    std::fill(s->rhs.begin(), s->rhs.end(), 0.0);
    for (size_t w = 0; w < s->sw.size(); w++) s->rhs[body_slot(s, int(w)) + 5 - n] = -1.0;
    solve_linear_system(s);
    for (size_t w = 0; w < s->sw.size(); w++)
      for (int k = 0; k < 6; k++) s->sw[w].Nref(k, 5 - n) = s->v_bodies[6 * w + k];
  }
  return 0;
}

// cilia_sim_main.cpp:205-340 for one step.  Returns 0, or -4 on a non-finite
// residual (divergence), -6 on Broyden non-convergence, -5 singular matrix.
int orc_sim_step(orc_sim *s, int nt, int *broyden_iters, int *gmres_iters) {
  s->total_gmres = 0;
  set_driving(s, nt);
  const double DL = s->DL();
  for (Swimmer &B : s->sw) sw_initial_guess(B, nt, s->p.num_euler_steps, DL);
  all_forces(s);
  sync_positions(s);
  compute_velocities(s);
  bool big = compute_errors(s, s->err, nt);
  if (big)
    for (Swimmer &B : s->sw)
      if (!sw_prepare_jacobian_inv(B, nt, s->p, DL)) return -5;
  s->iter = 0;
  while (big && s->iter < s->p.max_broyden_iter) {
    find_update(s);
    for (Swimmer &B : s->sw) sw_update(B, &s->upd[B.err0], DL);
    all_forces(s);
    sync_positions(s);
    compute_velocities(s);
    big = compute_errors(s, s->new_err, nt);
    for (double v : s->new_err)
      if (!std::isfinite(v)) {
        if (broyden_iters) *broyden_iters = s->iter + 1;
        return -4;
      }
    end_of_iter(s, big);
  }
  if (broyden_iters) *broyden_iters = s->iter;
  if (gmres_iters) *gmres_iters = int(s->total_gmres);
  return big ? -6 : 0;
}

int orc_sim_num_segments(const orc_sim *s) { return s->Ns; }
int orc_sim_num_blobs(const orc_sim *s) { return s->Nb; }
int orc_sim_nbroy(const orc_sim *s) { return s->nbroy; }

// Per-segment state: x[3Ns], q[4Ns] (s, v0, v1, v2), u[3Ns]; per-filament
// lambda[3(S-1)], tether[3], clamp[3]; per-swimmer body x[3], q[4], u[3].
void orc_sim_get_state(const orc_sim *s, double *seg_x, double *seg_q, double *seg_u,
                       double *fil_lam, double *body_xqu) {
  int g = 0, fg = 0;
  const int S = s->S;
  for (size_t n = 0; n < s->sw.size(); n++) {
    const Swimmer &B = s->sw[n];
    for (const Filament &F : B.fil) {
      for (int j = 0; j < S; j++, g++) {
        const Segment &sg = F.seg[j];
        if (seg_x) for (int k = 0; k < 3; k++) seg_x[3 * g + k] = sg.x[k];
        if (seg_q) { seg_q[4 * g] = sg.q.s; for (int k = 0; k < 3; k++) seg_q[4 * g + 1 + k] = sg.q.v[k]; }
        if (seg_u) for (int k = 0; k < 3; k++) seg_u[3 * g + k] = sg.u[k];
      }
      if (fil_lam) {
        double *o = &fil_lam[fg * (3 * (S - 1) + 6)];
        for (int k = 0; k < 3 * (S - 1); k++) o[k] = F.lam[k];
        for (int k = 0; k < 3; k++) { o[3 * (S - 1) + k] = F.teth[k]; o[3 * S + k] = F.clamp[k]; }
      }
      fg++;
    }
    if (body_xqu) {
      double *o = &body_xqu[10 * n];
      for (int k = 0; k < 3; k++) o[k] = B.x[k];
      o[3] = B.q.s;
      for (int k = 0; k < 3; k++) o[4 + k] = B.q.v[k];
      for (int k = 0; k < 3; k++) o[7 + k] = B.u[k];
    }
  }
}

// Backup record in the reference's order (no text): per swimmer its
// filaments (filament.cpp:1629-1648: per segment q, x, xm1, xm2, u, um1;
// tether, clamp and link multipliers with two histories), then its body
// (rigid_body.cpp:414-424: q, x, xm1, xm2, u, um1); swimmer.cpp:914-925.
}  // extern "C"
namespace {
template <typename S_, typename F>
void backup_walk(S_ *s, F &&io) {
  const int S = s->S;
  for (auto &B : s->sw) {
    for (auto &Fl : B.fil) {
      for (int j = 0; j < S; j++) {
        auto &g = Fl.seg[j];
        io(g.q.s);
        for (int k = 0; k < 3; k++) io(g.q.v[k]);
        for (int k = 0; k < 3; k++) io(g.x[k]);
        for (int k = 0; k < 3; k++) io(g.xm1[k]);
        for (int k = 0; k < 3; k++) io(g.xm2[k]);
        for (int k = 0; k < 3; k++) io(g.u[k]);
        for (int k = 0; k < 3; k++) io(g.um1[k]);
      }
      for (int k = 0; k < 3; k++) io(Fl.teth[k]);
      for (int k = 0; k < 3; k++) io(Fl.tethm1[k]);
      for (int k = 0; k < 3; k++) io(Fl.tethm2[k]);
      for (int k = 0; k < 3; k++) io(Fl.clamp[k]);
      for (int k = 0; k < 3; k++) io(Fl.clampm1[k]);
      for (int k = 0; k < 3; k++) io(Fl.clampm2[k]);
      for (int n = 0; n < S - 1; n++) {
        for (int k = 0; k < 3; k++) io(Fl.lam[3 * n + k]);
        for (int k = 0; k < 3; k++) io(Fl.lamm1[3 * n + k]);
        for (int k = 0; k < 3; k++) io(Fl.lamm2[3 * n + k]);
      }
    }
    io(B.q.s);
    for (int k = 0; k < 3; k++) io(B.q.v[k]);
    for (int k = 0; k < 3; k++) io(B.x[k]);
    for (int k = 0; k < 3; k++) io(B.xm1[k]);
    for (int k = 0; k < 3; k++) io(B.xm2[k]);
    for (int k = 0; k < 3; k++) io(B.u[k]);
    for (int k = 0; k < 3; k++) io(B.um1[k]);
  }
}
}  // namespace
extern "C" {

int orc_sim_backup_size(const orc_sim *s) {
  int n = 0;
  backup_walk(s, [&](const double &) { n++; });
  return n;
}

void orc_sim_get_backup(const orc_sim *s, double *out) {
  size_t i = 0;
  backup_walk(s, [&](const double &v) { out[i++] = v; });
}

// The reader (filament.cpp:84-163, rigid_body.cpp:28-56): q_m1 = q.
void orc_sim_set_backup(orc_sim *s, const double *in) {
  size_t i = 0;
  backup_walk(s, [&](double &v) { v = in[i++]; });
  for (auto &B : s->sw) {
    B.qm1 = B.q;
    for (auto &Fl : B.fil)
      for (auto &g : Fl.seg) g.qm1 = g.q;
  }
}

void orc_sim_get_mobility(const orc_sim *s, double *v_segs, double *v_bodies, double *f_blobs,
                          double *f_segs, double *body_ref66 /* Nref then KTKinv, per swimmer */) {
  if (v_segs) std::copy(s->vs.begin(), s->vs.end(), v_segs);
  if (v_bodies) std::copy(s->v_bodies.begin(), s->v_bodies.end(), v_bodies);
  if (f_blobs) std::copy(s->fb.begin(), s->fb.end(), f_blobs);
  if (f_segs) std::copy(s->fs.begin(), s->fs.end(), f_segs);
  if (body_ref66)
    for (size_t n = 0; n < s->sw.size(); n++) {
      std::copy(s->sw[n].Nref.d.begin(), s->sw[n].Nref.d.end(), body_ref66 + 72 * n);
      std::copy(s->sw[n].KTKinv.d.begin(), s->sw[n].KTKinv.d.end(), body_ref66 + 72 * n + 36);
    }
}

void orc_sim_get_broyden(const orc_sim *s, double *err) {
  if (err) std::copy(s->err.begin(), s->err.end(), err);
}

// Test hooks for pieces of the step (all on the current state).
int orc_sim_test_compute_velocities(orc_sim *s) {
  sync_positions(s);
  return compute_velocities(s);
}
void orc_sim_test_forces(orc_sim *s) { all_forces(s); }
int orc_sim_test_prepare_jacobian(orc_sim *s, int nt) {
  for (Swimmer &B : s->sw)
    if (!sw_prepare_jacobian_inv(B, nt, s->p, s->DL())) return -5;
  return 0;
}
// Approximate Jacobian (before inversion) of filament `fil` of swimmer 0,
// written to J[(6S)^2] row-major, and its inverse to Jinv.
int orc_sim_test_filament_jacobian(orc_sim *s, int fil, int nt, double *J, double *Jinv) {
  Filament &F = s->sw[0].fil[fil];
  const int n = 6 * s->S;
  Filament G = F;
  // rebuild the matrix without inverting: reuse the routine, then invert back
  if (!invert_approx_jacobian(G, nt, s->p, s->DL())) return -5;
  if (Jinv) std::copy(G.Jinv.d.begin(), G.Jinv.d.end(), Jinv);
  Mat A = G.Jinv;
  if (!invert(A)) return -5;
  if (J) std::copy(A.d.begin(), A.d.end(), J);
  (void)n;
  return 0;
}
void orc_sim_test_jacobian_inv_mult(orc_sim *s, const double *in, double *out) {
  for (const Swimmer &B : s->sw) sw_jacobian_inv_mult(B, &in[B.err0], &out[B.err0], s->DL());
}
void orc_sim_test_set_driving(orc_sim *s, int nt) { set_driving(s, nt); }
// Lie-group helpers exposed for tests (q as s, v0, v1, v2)
void orc_lie_exp(const double *u, double *q) {
  const Quat r = lie_exp(u);
  q[0] = r.s; q[1] = r.v[0]; q[2] = r.v[1]; q[3] = r.v[2];
}
void orc_lie_midpoint(const double *q1, const double *q2, double *q) {
  Quat a, b;
  a.s = q1[0]; b.s = q2[0];
  for (int k = 0; k < 3; k++) { a.v[k] = q1[1 + k]; b.v[k] = q2[1 + k]; }
  const Quat r = midpoint(a, b);
  q[0] = r.s; q[1] = r.v[0]; q[2] = r.v[1]; q[3] = r.v[2];
}
void orc_lie_rotmat(const double *q, double *R) {
  Quat a;
  a.s = q[0];
  for (int k = 0; k < 3; k++) a.v[k] = q[1 + k];
  const M3 m = rot_mat(a);
  for (int k = 0; k < 9; k++) R[k] = m.a[k];
}
void orc_lie_dexp(const double *u, const double *v, double *out) { dexp(out, u, v); }
void orc_lie_dexpinv(const double *u, const double *v, double *out) { dexpinv(out, u, v); }
void orc_lie_dexpinv_transpose(const double *u, const double *v, double *out) {
  dexpinv_transpose(out, u, v);
}
int orc_invert(double *A, int n) {
  Mat M(n, n);
  std::copy(A, A + n * n, M.d.begin());
  if (!invert(M)) return -5;
  std::copy(M.d.begin(), M.d.end(), A);
  return 0;
}

int orc_sim_test_errors(orc_sim *s, int nt, double *out) {
  std::vector<double> e(s->nbroy, 0.0);
  const bool big = compute_errors(s, e, nt);
  std::copy(e.begin(), e.end(), out);
  return big ? 1 : 0;
}
void orc_sim_test_initial_guess(orc_sim *s, int nt) {
  set_driving(s, nt);
  for (Swimmer &B : s->sw) sw_initial_guess(B, nt, s->p.num_euler_steps, s->DL());
}

void orc_test_filament_forces(int S, double DL, double kb, double kt, const double *q,
                              const double *u0, const double *lam, const double *teth,
                              const double *clamp, const double *st, double end_force,
                              const double *pert, double *f) {
  Filament F;
  F.seg.resize(S);
  for (int j = 0; j < S; j++) {
    F.seg[j].q.s = q[4 * j];
    for (int k = 0; k < 3; k++) F.seg[j].q.v[k] = q[4 * j + 1 + k];
  }
  for (int k = 0; k < 3; k++) {
    F.seg[0].u[k] = u0[k];
    F.teth[k] = teth[k];
    F.clamp[k] = clamp[k];
    F.st[k] = st[k];
  }
  F.lam.assign(lam, lam + 3 * (S - 1));
  if (pert)
    for (int k = 0; k < 6; k++) F.pert[k] = pert[k];
  F.f = f;
  internal_forces(F, DL, kb, kt, end_force, pert != nullptr);
}

void orc_sim_test_set_kinematics(orc_sim *s, const double *seg, const double *v_segs,
                                 const double *body, const double *v_bodies) {
  int g = 0;
  for (size_t n = 0; n < s->sw.size(); n++) {
    Swimmer &B = s->sw[n];
    for (Filament &F : B.fil)
      for (Segment &sg : F.seg) {
        const double *r = seg + 15 * g;
        for (int k = 0; k < 3; k++) {
          sg.x[k] = r[k]; sg.xm1[k] = r[3 + k]; sg.xm2[k] = r[6 + k];
          sg.u[k] = r[9 + k]; sg.um1[k] = r[12 + k];
        }
        g++;
      }
    const double *r = body + 15 * n;
    for (int k = 0; k < 3; k++) {
      B.x[k] = r[k]; B.xm1[k] = r[3 + k]; B.xm2[k] = r[6 + k];
      B.u[k] = r[9 + k]; B.um1[k] = r[12 + k];
    }
    for (int k = 0; k < 6; k++) s->v_bodies[6 * n + k] = v_bodies[6 * n + k];
  }
  std::copy(v_segs, v_segs + 6 * s->Ns, s->vs.begin());
}

int orc_sim_test_broyden_begin(orc_sim *s, int nt) {
  set_driving(s, nt);
  const double DL = s->DL();
  for (Swimmer &B : s->sw) sw_initial_guess(B, nt, s->p.num_euler_steps, DL);
  all_forces(s);
  sync_positions(s);
  compute_velocities(s);
  const bool big = compute_errors(s, s->err, nt);
  for (Swimmer &B : s->sw)
    if (!sw_prepare_jacobian_inv(B, nt, s->p, DL)) return -5;
  s->iter = 0;
  return big ? 1 : 0;
}

int orc_sim_test_broyden_iterate(orc_sim *s, int nt, double *dx, double *e_old, double *e_new) {
  const double DL = s->DL();
  find_update(s);
  std::copy(s->upd.begin(), s->upd.end(), dx);
  for (Swimmer &B : s->sw) sw_update(B, &s->upd[B.err0], DL);
  all_forces(s);
  sync_positions(s);
  compute_velocities(s);
  const bool big = compute_errors(s, s->new_err, nt);
  std::copy(s->err.begin(), s->err.end(), e_old);
  std::copy(s->new_err.begin(), s->new_err.end(), e_new);
  end_of_iter(s, true);
  return big ? 1 : 0;
}

int orc_sim_test_broyden_columns(const orc_sim *s, double *C, double *D) {
  std::copy(s->C.d.begin(), s->C.d.end(), C);
  std::copy(s->D.d.begin(), s->D.d.end(), D);
  return s->iter;
}

}  // extern "C"
