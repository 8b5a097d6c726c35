// cilia.cpp -- oracle (test infrastructure only): FP64 CPU restatement of the
// reference's prescribed-cilia time step (CILIA_TYPE 3, GMRES solver):
//   driver loop        cilia_sim_main.cpp:205-232 (explicit body update),
//                      320-330 (end_of_step)
//   swimmer            swimmer.cpp:240-275 (initial_guess), 927-1027
//                      (make_precon_mat)
//   filament           filament.cpp:197-282 (beat frame and phase), 465-815
//                      (beat shapes), 825-893 (initial_guess), 1020-1031
//                      (end_of_step)
//   rigid body         rigid_body.cpp:334-352 (blob positions), 436-450
//                      (prescribed velocities)
//   mobility solver    mobility_solver.cpp:36-60 (generalised-force table),
//                      196-359 (assemble_rhs), 427-650 (apply_preconditioner),
//                      733-904 (system_matrix_mult), 970-1354
//                      (solve_linear_system, right preconditioning),
//                      1358-1366 (compute_velocities)
//   pair blocks        cuda_functions.cu:246-863, translation parts only
//                      (the rotation terms are compiled out, :345,646,808)
// Build-specific choices (kept identical on the device):
//  * the phase block of the preconditioner is negated and divided by
//    6 pi mu a_s once per swimmer for that swimmer's own filaments (the
//    reference applies both to every swimmer's phase entries inside its
//    swimmer loop, mobility_solver.cpp:518,622; that only changes the
//    Krylov path for NSWIM > 1, never the solution);
//  * swimmers either all couple hydrodynamically (the reference) or are
//    independent (interact = 0).
#include "oracle_cilia.h"
#include "oracle.h"
#include "lie.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

using namespace orc;

namespace {

// Fourier coefficients of the fitted beats (filament.cpp:471-549): rows are
// the cosine modes 0..3 (A) or sine modes 1..3 (B), columns the powers s,
// s^2, s^3 of the arc-length parameter.  Transposed storage as in the
// reference; the zero modes carry the factor 1/2.
struct Fourier {
  double Ax[4][3], Ay[4][3], Bx[3][3], By[3][3];
};

Fourier fourier_table(int beat) {
  Fourier F;
  if (beat == 1) {  // Fulford & Blake (Ito, Omori & Ishikawa 2019, Table 1)
    const double Ay[4][3] = {{-3.3547e-01, 4.0369e-01, 1.0362e-01},
                             {4.0318e-01, -1.5553e+00, 7.3455e-01},
                             {-9.9513e-02, 3.2829e-02, -1.2106e-01},
                             {8.1046e-02, -3.0982e-01, 1.4568e-01}};
    const double Ax[4][3] = {{9.7204e-01, -2.8315e-01, 4.9243e-02},
                             {-1.8466e-02, -1.2926e-01, 2.6981e-01},
                             {1.6209e-01, -3.4983e-01, 1.9082e-01},
                             {1.0259e-02, 3.5907e-02, -6.8736e-02}};
    const double By[3][3] = {{2.9136e-01, 1.0721e+00, -1.0433e+00},
                             {6.1554e-03, 3.2521e-01, -2.8315e-01},
                             {-6.0528e-02, 2.3185e-01, -2.0108e-01}};
    const double Bx[3][3] = {{1.9697e-01, -5.1193e-01, 3.4778e-01},
                             {-5.1295e-02, 4.3396e-01, -3.3547e-01},
                             {1.2311e-02, 1.4157e-01, -1.1695e-01}};
    std::memcpy(F.Ax, Ax, sizeof Ax); std::memcpy(F.Ay, Ay, sizeof Ay);
    std::memcpy(F.Bx, Bx, sizeof Bx); std::memcpy(F.By, By, sizeof By);
  } else {  // coral larvae
    const double Ay[4][3] = {{-2.766076e-01, 6.368662e-01, -8.513947e-02},
                             {3.091205e-01, -1.993712e+00, 1.091847e+00},
                             {-2.952120e-01, 7.189021e-01, -6.328589e-01},
                             {-9.868018e-02, 3.312812e-01, -2.578206e-01}};
    const double Ax[4][3] = {{9.034276e-01, -6.634681e-01, 2.016123e-01},
                             {-2.468656e-01, 1.215869e+00, -5.867236e-01},
                             {3.075836e-01, -6.966322e-01, 4.554729e-01},
                             {2.378489e-02, -1.355582e-01, 7.315831e-02}};
    const double By[3][3] = {{9.165546e-01, 8.848315e-02, -4.838832e-01},
                             {-2.946211e-01, 1.057964e+00, -6.335721e-01},
                             {6.122744e-02, -2.319094e-01, 1.906292e-01}};
    const double Bx[3][3] = {{1.336807e-01, -6.807852e-01, 6.008592e-01},
                             {8.146824e-02, 3.472676e-01, -2.744220e-01},
                             {3.615272e-02, 8.619119e-02, -6.122992e-02}};
    std::memcpy(F.Ax, Ax, sizeof Ax); std::memcpy(F.Ay, Ay, sizeof Ay);
    std::memcpy(F.Bx, Bx, sizeof Bx); std::memcpy(F.By, By, sizeof By);
  }
  return F;
}

// row vector (c A + s B) times the power vector w, for one component
double fourier_eval(const double (*A)[3], const double (*B)[3], const double *c,
                    const double *s, const double *w) {
  double out = 0.0;
  for (int j = 0; j < 3; j++) {
    double r = 0.0;
    for (int k = 0; k < 4; k++) r += c[k] * A[k][j];
    double t = 0.0;
    for (int k = 0; k < 3; k++) t += s[k] * B[k][j];
    out += (r + t) * w[j];
  }
  return out;
}

struct CFil {
  int sw = 0;
  double ref[3] = {0, 0, 0};
  double phase = 0, omega0 = 0, speed = 0;
  Quat q, qm1;
  std::vector<double> K;  // vel_dir [3 S]
};

struct CSw {
  int nfil = 0, nblob = 0, seg0 = 0, blob0 = 0, fil0 = 0;
  double x[3] = {0, 0, 0}, u[3] = {0, 0, 0};
  Quat q, qm1;
  std::vector<double> blob_ref;
  Mat KI;                  // KTMinvK_inv (6x6)
  double vp[6] = {0, 0, 0, 0, 0, 0};  // prescribed (V, W)
};

}  // namespace

struct orc_cilia {
  orc_cilia_params p;
  int S = 0, Nsw = 0, Ns = 0, Nb = 0, Nf = 0, nsys = 0, nthreads = 1;
  double DL = 0;
  Fourier F;
  std::vector<CSw> sw;
  std::vector<CFil> fil;
  std::vector<double> gref;
  std::vector<double> xs, xb;                  // positions (3 Ns, 3 Nb)
  std::vector<double> f_segs, f_blobs, v_bodies;  // solution arrays
  std::vector<double> rhs;
  int last_iters = 0;

  int off_b() const { return 3 * Ns; }
  int off_c() const { return 3 * Ns + 3 * Nb; }
  int off_p() const { return off_c() + (p.fixed_body ? 0 : 6 * Nsw); }

  // ---- beat shapes (filament.cpp:465-815) --------------------------------
  void trig(double phase, double *c, double *s) const {
    c[0] = 1.0;
    for (int n = 1; n < 4; n++) { c[n] = std::cos(n * phase); s[n - 1] = std::sin(n * phase); }
  }
  void fitted_tangent(double phase, double sv, double &tx, double &ty) const {
    double c[4], s[3];
    trig(phase, c, s);
    const double w[3] = {1.0, 2.0 * sv, 3.0 * sv * sv};
    tx = fourier_eval(F.Ax, F.Bx, c, s, w);
    ty = fourier_eval(F.Ay, F.By, c, s, w);
  }
  void fitted_shape(double phase, double sv, double *pos) const {
    double c[4], s[3];
    trig(phase, c, s);
    const double w[3] = {sv, sv * sv, sv * sv * sv};
    pos[0] = fourier_eval(F.Ax, F.Bx, c, s, w);
    pos[1] = fourier_eval(F.Ay, F.By, c, s, w);
    pos[2] = 0.0;
  }
  // d(shape)/d(phase): cosine modes -> -n sin(n phase), sine modes -> n cos(n phase)
  void fitted_dir(double phase, double sv, double *dir) const {
    double c[4], s[3];
    c[0] = 0.0;
    for (int n = 1; n < 4; n++) { c[n] = -n * std::sin(n * phase); s[n - 1] = n * std::cos(n * phase); }
    const double w[3] = {sv, sv * sv, sv * sv * sv};
    dir[0] = fourier_eval(F.Ax, F.Bx, c, s, w);
    dir[1] = fourier_eval(F.Ay, F.By, c, s, w);
    dir[2] = 0.0;
  }
  double tan_norm(double phase, double sv) const {
    double tx, ty;
    fitted_tangent(phase, sv, tx, ty);
    return std::sqrt(tx * tx + ty * ty);
  }
  // filament.cpp:618-666: trapezoids on a fixed mesh of 10 NSEG intervals
  double curve_length(double phase, double sv) const {
    double length = 0.0;
    if (sv > 0) {
      const int ntr = 10 * S;
      const double dl = 1.0 / double(ntr);
      double f0 = tan_norm(phase, 0.0);
      const double fl = std::floor(sv / dl);
      for (int n = 1; n <= fl; n++) {
        const double f1 = tan_norm(phase, n * dl);
        length += 0.5 * dl * (f0 + f1);
        f0 = f1;
      }
      const double cl = std::ceil(sv / dl);
      if (cl > fl) {
        const double f1 = tan_norm(phase, cl * dl);
        length += ((sv / dl - fl) / (cl - fl)) * 0.5 * dl * (f0 + f1);
      }
    }
    return length;
  }
  // filament.cpp:668-715: arc-length parameters by bisection
  std::vector<double> fitted_s(double phase) const {
    std::vector<double> su(S);
    su[0] = 0.0;
    su[S - 1] = 1.0;
    const double total = curve_length(phase, 1.0);
    for (int n = 1; n < S - 1; n++) {
      const double target = double(n) / double(S - 1);
      double lo = su[n - 1], hi = 1.0;
      double est = 0.5 * (lo + hi);
      double frac = curve_length(phase, est) / total;
      while (std::fabs(frac - target) > 0.1 / double(S)) {
        if (frac > target) hi = est; else lo = est;
        est = 0.5 * (lo + hi);
        frac = curve_length(phase, est) / total;
      }
      su[n] = est;
    }
    return su;
  }
  // filament.cpp:719-763
  double bab_angle(double phase, double sv) const {
    const double wR = p.recovery_window * 2.0 * kPi;
    const double wE = p.effective_stroke * 2.0 * kPi;
    const double phi_tr = wE + sv * (2.0 * kPi - wE - wR);
    double sp = phase - sv * 2.0 * kPi * p.zero_velocity_avoidance;
    sp -= 2.0 * kPi * std::floor(0.5 * sp / kPi);
    double delta;
    if (sp < wE) {
      const double t = kPi * (wE - sp) / wE;
      delta = t - std::sin(t) * std::cos(t);
    } else if (sp < phi_tr) {
      delta = 0.0;
    } else if (sp < phi_tr + wR) {
      const double t = kPi * (sp - phi_tr) / wR;
      delta = t - std::sin(t) * std::cos(t);
    } else {
      delta = kPi;
    }
    const double th = std::acos(0.5 * p.amplitude);
    return th + (kPi - 2.0 * th) * delta / kPi;
  }
  void bab_tangent(double phase, double sv, double *t) const {
    const double a = bab_angle(phase, sv);
    t[0] = std::sin(a); t[1] = std::cos(a); t[2] = 0.0;
  }
  // filament.cpp:775-813
  void bab_tangent_deriv(double phase, double sv, double *k) const {
    const double th = std::acos(0.5 * p.amplitude);
    const double fac = 2.0 * (kPi - 2.0 * th);
    const double wR = p.recovery_window * 2.0 * kPi;
    const double wE = p.effective_stroke * 2.0 * kPi;
    const double phi_tr = wE + sv * (2.0 * kPi - wE - wR);
    double sp = phase - sv * 2.0 * kPi * p.zero_velocity_avoidance;
    sp -= 2.0 * kPi * std::floor(0.5 * sp / kPi);
    double dd;
    if (sp < wE) {
      const double t = kPi * (wE - sp) / wE;
      dd = std::sin(t);
      dd *= -dd / wE;
    } else if (sp < phi_tr) {
      dd = 0.0;
    } else if (sp < phi_tr + wR) {
      const double t = kPi * (sp - phi_tr) / wR;
      dd = std::sin(t);
      dd *= dd / wR;
    } else {
      dd = 0.0;
    }
    const double a = bab_angle(phase, sv);
    k[0] = fac * dd * std::cos(a);
    k[1] = -fac * dd * std::sin(a);
    k[2] = 0.0;
  }

  // filament.cpp:825-893 (prescribed branch): shape and velocity direction
  // of filament f from its base position and frame
  void filament_guess(int f, const double *base, const double *u) {
    CFil &c = fil[f];
    double *x = &xs[3 * size_t(f) * S];
    x[0] = base[0]; x[1] = base[1]; x[2] = base[2];
    c.q = qmul(lie_exp(u), c.qm1);
    c.speed = c.omega0;
    const M3 R = rot_mat(c.q);
    c.K.assign(3 * S, 0.0);
    if (p.beat == 0) {
      double t1[3], t2[3], k1[3], k2[3], tmp[3];
      bab_tangent(c.phase, 0.0, tmp); mv3(R, tmp, t1);
      bab_tangent_deriv(c.phase, 0.0, tmp); mv3(R, tmp, k1);
      for (int n = 1; n < S; n++) {
        const double sv = double(n) / double(S - 1);
        bab_tangent(c.phase, sv, tmp); mv3(R, tmp, t2);
        for (int k = 0; k < 3; k++) x[3 * n + k] = x[3 * (n - 1) + k] + 0.5 * DL * (t1[k] + t2[k]);
        std::memcpy(t1, t2, sizeof t1);
        bab_tangent_deriv(c.phase, sv, tmp); mv3(R, tmp, k2);
        for (int k = 0; k < 3; k++) c.K[3 * n + k] = c.K[3 * (n - 1) + k] + 0.5 * DL * (k1[k] + k2[k]);
        std::memcpy(k1, k2, sizeof k1);
      }
    } else {
      const double L = DL * (S - 1);
      const std::vector<double> su = fitted_s(c.phase);
      const M3 LR = L * R;
      for (int n = 1; n < S; n++) {
        double v[3], d[3];
        fitted_shape(c.phase, su[n], v);
        mv3(LR, v, d);
        for (int k = 0; k < 3; k++) x[3 * n + k] = x[k] + d[k];
        fitted_dir(c.phase, su[n], v);
        mv3(LR, v, d);
        for (int k = 0; k < 3; k++) c.K[3 * n + k] = d[k];
      }
    }
  }

  // swimmer.cpp:929-1027
  void make_precon_mat(int n) {
    CSw &B = sw[n];
    B.KI = Mat(6, 6);
    if (p.fixed_body) return;
    const double sf = 6.0 * kPi * p.mu * p.a_seg, bf = 6.0 * kPi * p.mu * p.a_blob;
    const double eye = double(B.nfil * S) * sf + double(B.nblob) * bf;
    Mat &K = B.KI;
    K(0, 0) = K(1, 1) = K(2, 2) = eye;
    for (int i = 0; i < B.nfil * S; i++) {
      const double *x = &xs[3 * size_t(B.seg0 + i)];
      const double d[3] = {x[0] - B.x[0], x[1] - B.x[1], x[2] - B.x[2]};
      const M3 rc = rcross(d);
      K.add3(0, 3, sf * rc);
      K.add3(3, 3, -1.0 * ((sf * rc) * rc));
    }
    const M3 Q = rot_mat(B.q);
    for (int m = 0; m < B.nblob; m++) {
      double d[3];
      mv3(Q, &B.blob_ref[3 * m], d);
      const M3 rc = rcross(d);
      K.add3(0, 3, bf * rc);
      K.add3(3, 3, -1.0 * ((bf * rc) * rc));
    }
    K.set3(3, 0, -1.0 * K.get3(0, 3));
    if (p.dynamic_phase) {
      Mat C(6, 6);
      for (int i = 0; i < B.nfil; i++) {
        double kn, v1[3], v2[3];
        fil_vectors(B.fil0 + i, kn, v1, v2);
        for (int a = 0; a < 3; a++)
          for (int b = 0; b < 3; b++) {
            C(a, b) += v1[a] * v1[b] / kn;
            C(3 + a, b) += v2[a] * v1[b] / kn;
            C(3 + a, 3 + b) += v2[a] * v2[b] / kn;
          }
      }
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) C(a, 3 + b) = C(3 + b, a);
      for (int k = 0; k < 36; k++) K.d[k] -= sf * C.d[k];
    }
    invert(K);
  }

  // |K|^2, sum K and sum (x - X) x K of one filament (swimmer.cpp:974-1008)
  void fil_vectors(int f, double &kn, double *v1, double *v2) const {
    const CFil &c = fil[f];
    const CSw &B = sw[c.sw];
    kn = 0.0;
    v1[0] = v1[1] = v1[2] = 0.0;
    v2[0] = v2[1] = v2[2] = 0.0;
    for (int m = 0; m < S; m++) {
      const double *k = &c.K[3 * m];
      const double *x = &xs[3 * (size_t(f) * S + m)];
      const double d[3] = {x[0] - B.x[0], x[1] - B.x[1], x[2] - B.x[2]};
      double cr[3];
      cross(d, k, cr);
      for (int a = 0; a < 3; a++) { v1[a] += k[a]; v2[a] += cr[a]; }
      kn += k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
    }
  }

  void blob_positions() {
    for (const CSw &B : sw) {
      const M3 R = rot_mat(B.q);
      for (int m = 0; m < B.nblob; m++) {
        double r[3];
        mv3(R, &B.blob_ref[3 * m], r);
        for (int k = 0; k < 3; k++) xb[3 * size_t(B.blob0 + m) + k] = B.x[k] + r[k];
      }
    }
  }

  // ---- pair blocks, translations only ------------------------------------
  // vs = M_ss fs + M_sb fb, vb = M_bs fs + M_bb fb
  void mobility(const double *fs3, const double *fb3, double *vs3, double *vb3) {
    std::vector<double> F6(6 * size_t(Ns), 0.0), V6(6 * size_t(Ns), 0.0), vb(3 * size_t(Nb), 0.0),
        vb2(3 * size_t(Nb), 0.0);
    for (int i = 0; i < Ns; i++)
      for (int k = 0; k < 3; k++) F6[6 * size_t(i) + k] = fs3[3 * size_t(i) + k];
    auto group = [&](int s0, int ns, int b0, int nb) {
      const double *X = &xs[3 * size_t(s0)];
      const double *Xb = &xb[3 * size_t(b0)];
      if (p.wall) {
        // cuda_functions.cu:8-242 (zero torques: the PRESCRIBED_CILIA rows)
        orc_Mss_wall(&V6[6 * size_t(s0)], &F6[6 * size_t(s0)], X, ns, 0, ns, p.a_seg, p.mu,
                     nthreads);
        return;
      }
      orc_Mss(&V6[6 * size_t(s0)], &F6[6 * size_t(s0)], X, ns, 0, ns, p.a_seg, p.mu, nthreads);
      if (nb > 0) {
        orc_Msb(&V6[6 * size_t(s0)], fb3 + 3 * size_t(b0), X, ns, Xb, nb, 0, ns, p.a_seg,
                p.a_blob, p.mu, nthreads);
        orc_Mbb(&vb[3 * size_t(b0)], fb3 + 3 * size_t(b0), Xb, nb, 0, nb, p.a_blob, p.mu,
                nthreads);
        orc_Mbs(&vb2[3 * size_t(b0)], &F6[6 * size_t(s0)], Xb, nb, X, ns, 0, nb, p.a_seg,
                p.a_blob, p.mu, nthreads);
      }
    };
    if (p.interact || Nsw == 1) {
      group(0, Ns, 0, Nb);
    } else {
      for (const CSw &B : sw) group(B.seg0, B.nfil * S, B.blob0, B.nblob);
    }
    for (int i = 0; i < Ns; i++)
      for (int k = 0; k < 3; k++) vs3[3 * size_t(i) + k] = V6[6 * size_t(i) + k];
    for (size_t e = 0; e < vb.size(); e++) vb3[e] = vb[e] + vb2[e];
  }

  // mobility_solver.cpp:733-904
  void sys_mult(const double *in, double *out) {
    std::fill(out, out + nsys, 0.0);
    const int ob = off_b(), oc = off_c(), op = off_p();
    if (!p.fixed_body) {
      for (int n = 0; n < Nsw; n++) {
        const CSw &B = sw[n];
        const double *Y = in + oc + 6 * n;
        double *T = out + oc + 6 * n;
        for (int i = B.seg0; i < B.seg0 + B.nfil * S; i++) {
          double *o = out + 3 * size_t(i);
          const double *x = &xs[3 * size_t(i)];
          const double d[3] = {x[0] - B.x[0], x[1] - B.x[1], x[2] - B.x[2]};
          o[0] -= Y[0]; o[1] -= Y[1]; o[2] -= Y[2];
          o[0] -= Y[4] * d[2] - Y[5] * d[1];
          o[1] -= Y[5] * d[0] - Y[3] * d[2];
          o[2] -= Y[3] * d[1] - Y[4] * d[0];
          const double *f = in + 3 * size_t(i);
          T[0] -= f[0]; T[1] -= f[1]; T[2] -= f[2];
          T[3] -= d[1] * f[2] - d[2] * f[1];
          T[4] -= d[2] * f[0] - d[0] * f[2];
          T[5] -= d[0] * f[1] - d[1] * f[0];
        }
        const M3 R = rot_mat(B.q);
        for (int m = 0; m < B.nblob; m++) {
          double *o = out + ob + 3 * size_t(B.blob0 + m);
          double d[3];
          mv3(R, &B.blob_ref[3 * m], d);
          o[0] -= Y[0]; o[1] -= Y[1]; o[2] -= Y[2];
          o[0] -= Y[4] * d[2] - Y[5] * d[1];
          o[1] -= Y[5] * d[0] - Y[3] * d[2];
          o[2] -= Y[3] * d[1] - Y[4] * d[0];
          const double *f = in + ob + 3 * size_t(B.blob0 + m);
          T[0] -= f[0]; T[1] -= f[1]; T[2] -= f[2];
          T[3] -= d[1] * f[2] - d[2] * f[1];
          T[4] -= d[2] * f[0] - d[0] * f[2];
          T[5] -= d[0] * f[1] - d[1] * f[0];
        }
      }
    }
    if (p.dynamic_phase) {
      for (int f = 0; f < Nf; f++) {
        const double phi_dot = in[op + f];
        double *o = out + 3 * size_t(f) * S;
        const double *z = in + 3 * size_t(f) * S;
        for (int k = 0; k < 3 * S; k++) {
          o[k] -= phi_dot * fil[f].K[k];
          out[op + f] -= z[k] * fil[f].K[k];
        }
      }
    }
    std::vector<double> vs(3 * size_t(Ns)), vb(3 * size_t(Nb));
    mobility(in, in + ob, vs.data(), vb.data());
    for (int e = 0; e < 3 * Nb; e++) out[ob + e] += vb[e];
    for (int e = 0; e < 3 * Ns; e++) out[e] += vs[e];
  }

  // mobility_solver.cpp:427-650 (right preconditioner)
  void precond(const double *in, double *out) {
    std::copy(in, in + nsys, out);
    const double sf = 6.0 * kPi * p.mu * p.a_seg, bf = 6.0 * kPi * p.mu * p.a_blob;
    const int ob = off_b(), oc = off_c(), op = off_p();
    for (int e = 0; e < 3 * Ns; e++) out[e] *= sf;
    for (int e = 0; e < 3 * Nb; e++) out[ob + e] *= bf;
    if (p.fixed_body) {
      if (!p.dynamic_phase) return;
      for (int f = 0; f < Nf; f++) out[op + f] *= -1.0;
      for (int f = 0; f < Nf; f++) {
        double t = 0.0, kn = 0.0;
        const std::vector<double> &K = fil[f].K;
        for (int k = 0; k < 3 * S; k++) {
          t += K[k] * out[3 * size_t(f) * S + k];
          kn += K[k] * K[k];
        }
        out[op + f] -= t;
        out[op + f] /= kn * sf;
        for (int k = 0; k < 3 * S; k++) out[3 * size_t(f) * S + k] += sf * out[op + f] * K[k];
      }
      return;
    }
    for (int e = 0; e < 6 * Nsw; e++) out[oc + e] *= -1.0;
    for (int n = 0; n < Nsw; n++) {
      const CSw &B = sw[n];
      double *Y = out + oc + 6 * n;
      for (int i = B.seg0; i < B.seg0 + B.nfil * S; i++) {
        const double *f = out + 3 * size_t(i);
        const double *x = &xs[3 * size_t(i)];
        const double d[3] = {x[0] - B.x[0], x[1] - B.x[1], x[2] - B.x[2]};
        Y[0] -= f[0]; Y[1] -= f[1]; Y[2] -= f[2];
        Y[3] -= d[1] * f[2] - d[2] * f[1];
        Y[4] -= d[2] * f[0] - d[0] * f[2];
        Y[5] -= d[0] * f[1] - d[1] * f[0];
      }
      const M3 Q = rot_mat(B.q);
      for (int m = 0; m < B.nblob; m++) {
        const double *f = out + ob + 3 * size_t(B.blob0 + m);
        double d[3];
        mv3(Q, &B.blob_ref[3 * m], d);
        Y[0] -= f[0]; Y[1] -= f[1]; Y[2] -= f[2];
        Y[3] -= d[1] * f[2] - d[2] * f[1];
        Y[4] -= d[2] * f[0] - d[0] * f[2];
        Y[5] -= d[0] * f[1] - d[1] * f[0];
      }
      if (p.dynamic_phase) {
        for (int i = 0; i < B.nfil; i++) out[op + B.fil0 + i] *= -1.0;
        for (int i = 0; i < B.nfil; i++) {
          const int f = B.fil0 + i;
          double kn, v1[3], v2[3], t = 0.0;
          fil_vectors(f, kn, v1, v2);
          const double *z = out + 3 * size_t(f) * S;
          for (int k = 0; k < S; k++) {
            const double *K = &fil[f].K[3 * k];
            t += K[0] * z[3 * k] + K[1] * z[3 * k + 1] + K[2] * z[3 * k + 2];
          }
          double &ph = out[op + f];
          ph -= t;
          ph /= kn;
          for (int a = 0; a < 3; a++) { Y[a] -= ph * v1[a]; Y[3 + a] -= ph * v2[a]; }
        }
      }
      double b[6];
      for (int a = 0; a < 6; a++) {
        double s = 0.0;
        for (int c = 0; c < 6; c++) s += B.KI(a, c) * Y[c];
        b[a] = s;
      }
      std::copy(b, b + 6, Y);
      if (p.dynamic_phase) {
        for (int i = 0; i < B.nfil; i++) out[op + B.fil0 + i] /= sf;
        for (int i = 0; i < B.nfil; i++) {
          const int f = B.fil0 + i;
          double kn, v1[3], v2[3];
          fil_vectors(f, kn, v1, v2);
          out[op + f] -= (v1[0] * Y[0] + v1[1] * Y[1] + v1[2] * Y[2] + v2[0] * Y[3] +
                          v2[1] * Y[4] + v2[2] * Y[5]) / kn;
        }
      }
      for (int i = 0; i < B.nfil; i++) {
        const int f = B.fil0 + i;
        for (int k = 0; k < S; k++) {
          const size_t id = 3 * (size_t(f) * S + k);
          const double *x = &xs[id];
          const double d[3] = {x[0] - B.x[0], x[1] - B.x[1], x[2] - B.x[2]};
          out[id] += sf * (Y[0] + Y[4] * d[2] - Y[5] * d[1]);
          out[id + 1] += sf * (Y[1] + Y[5] * d[0] - Y[3] * d[2]);
          out[id + 2] += sf * (Y[2] + Y[3] * d[1] - Y[4] * d[0]);
          if (p.dynamic_phase) {
            const double ph = out[op + f];
            for (int a = 0; a < 3; a++) out[id + a] += sf * ph * fil[f].K[3 * k + a];
          }
        }
      }
      for (int m = 0; m < B.nblob; m++) {
        const size_t id = ob + 3 * size_t(B.blob0 + m);
        double d[3];
        mv3(Q, &B.blob_ref[3 * m], d);
        out[id] += bf * (Y[0] + Y[4] * d[2] - Y[5] * d[1]);
        out[id + 1] += bf * (Y[1] + Y[5] * d[0] - Y[3] * d[2]);
        out[id + 2] += bf * (Y[2] + Y[3] * d[1] - Y[4] * d[0]);
      }
    }
  }

  // mobility_solver.cpp:223-306: cubic Hermite interpolation of the table
  double interp(double phase, double omega0) const {
    const int N = int(gref.size());
    double pi = 0.5 * phase / kPi;
    pi -= std::floor(pi);
    pi *= N;
    double q;
    const int lo = int(pi);
    if (lo == N) {
      q = gref[0];
    } else {
      const int up = (lo == N - 1) ? 0 : lo + 1;
      const double ql = gref[lo];
      const double qpl = (lo == 0) ? (gref[1] - gref[0]) : 0.5 * (gref[up] - gref[lo - 1]);
      const double qu = gref[up];
      const double qpu = (up == 0) ? (gref[0] - gref[N - 1])
                                   : 0.5 * (gref[(up == N - 1) ? 0 : up + 1] - gref[lo]);
      const double t = pi - lo, t2 = t * t, t3 = t2 * t;
      q = ql * (2.0 * t3 - 3.0 * t2 + 1.0) + qpl * (t3 - 2.0 * t2 + t) + qu * (3.0 * t2 - 2.0 * t3) +
          qpu * (t3 - t2);
    }
    return q * 0.5 * omega0 / kPi;
  }

  // mobility_solver.cpp:196-359 (prescribed branch)
  void assemble_rhs(int nt) {
    (void)nt;
    rhs.assign(nsys, 0.0);
    const int ob = off_b(), op = off_p();
    for (int f = 0; f < Nf; f++) {
      if (p.dynamic_phase) {
        rhs[op + f] = -interp(fil[f].phase, fil[f].omega0);
      } else {
        for (int m = 0; m < 3 * S; m++) rhs[3 * size_t(f) * S + m] = fil[f].speed * fil[f].K[m];
      }
    }
    if (p.fixed_body) {
      for (int n = 0; n < Nsw; n++) {
        const CSw &B = sw[n];
        for (int k = 0; k < 6; k++) v_bodies[6 * n + k] = B.vp[k];
        const double *Y = &v_bodies[6 * n];
        const M3 R = rot_mat(B.q);
        for (int m = 0; m < B.nblob; m++) {
          double r[3];
          mv3(R, &B.blob_ref[3 * m], r);
          double *o = &rhs[ob + 3 * size_t(B.blob0 + m)];
          o[0] += Y[0] + Y[4] * r[2] - Y[5] * r[1];
          o[1] += Y[1] + Y[5] * r[0] - Y[3] * r[2];
          o[2] += Y[2] + Y[3] * r[1] - Y[4] * r[0];
        }
        for (int i = B.seg0; i < B.seg0 + B.nfil * S; i++) {
          const double *x = &xs[3 * size_t(i)];
          const double r[3] = {x[0] - B.x[0], x[1] - B.x[1], x[2] - B.x[2]};
          double *o = &rhs[3 * size_t(i)];
          o[0] += Y[0] + Y[4] * r[2] - Y[5] * r[1];
          o[1] += Y[1] + Y[5] * r[0] - Y[3] * r[2];
          o[2] += Y[2] + Y[3] * r[1] - Y[4] * r[0];
        }
      }
    }
  }

  void unpack(const std::vector<double> &x) {
    const int ob = off_b(), oc = off_c(), op = off_p();
    std::copy(x.begin(), x.begin() + 3 * size_t(Ns), f_segs.begin());
    std::copy(x.begin() + ob, x.begin() + ob + 3 * size_t(Nb), f_blobs.begin());
    if (!p.fixed_body) std::copy(x.begin() + oc, x.begin() + oc + 6 * size_t(Nsw), v_bodies.begin());
    if (p.dynamic_phase)
      for (int f = 0; f < Nf; f++) fil[f].speed = x[op + f];
  }

  // mobility_solver.cpp:970-1354, right preconditioning, modified Gram-Schmidt
  int solve() {
    const int n = nsys;
    auto nrm = [&](const double *v) {
      double s = 0.0;
      for (int e = 0; e < n; e++) s += v[e] * v[e];
      return std::sqrt(s);
    };
    const double nrhs = nrm(rhs.data());
    std::vector<double> soln(n, 0.0);
    if (nrhs == 0.0) {
      unpack(soln);
      if (p.dynamic_phase)
        for (int f = 0; f < Nf; f++) fil[f].speed = 0.0;
      return 0;
    }
    const int ob = off_b(), oc = off_c(), op = off_p();
    std::copy(f_segs.begin(), f_segs.end(), soln.begin());
    std::copy(f_blobs.begin(), f_blobs.end(), soln.begin() + ob);
    if (!p.fixed_body) std::copy(v_bodies.begin(), v_bodies.end(), soln.begin() + oc);
    if (p.dynamic_phase)
      for (int f = 0; f < Nf; f++) soln[op + f] = fil[f].speed;
    const int maxit = std::min(n, p.gmres_max_iter);
    std::vector<double> Q(size_t(n) * (maxit + 1), 0.0), H(size_t(maxit + 1) * (maxit + 1), 0.0),
        beta(maxit + 1, 0.0), cs(maxit, 0.0), sn(maxit, 0.0), t1(n), t2(n);
    auto col = [&](int j) { return &Q[size_t(j) * n]; };
    auto h = [&](int i, int j) -> double & { return H[size_t(i) * (maxit + 1) + j]; };
    sys_mult(soln.data(), t1.data());
    for (int e = 0; e < n; e++) col(0)[e] = rhs[e] - t1[e];
    beta[0] = nrm(col(0));
    if (std::fabs(beta[0]) <= p.gmres_tol * nrhs) return 0;
    if (std::fabs(beta[0]) > nrhs) {
      std::fill(soln.begin(), soln.end(), 0.0);
      beta[0] = nrhs;
      std::copy(rhs.begin(), rhs.end(), col(0));
    }
    for (int e = 0; e < n; e++) col(0)[e] /= beta[0];
    for (int it = 1; it <= maxit; it++) {
      precond(col(it - 1), t1.data());
      sys_mult(t1.data(), col(it));
      double *w = col(it);
      for (int i = 0; i < it; i++) {
        double d = 0.0;
        const double *qi = col(i);
        for (int e = 0; e < n; e++) d += w[e] * qi[e];
        h(i, it - 1) = d;
        for (int e = 0; e < n; e++) w[e] -= d * qi[e];
      }
      const double qn = nrm(w);
      h(it, it - 1) = qn;
      for (int e = 0; e < n; e++) w[e] /= qn;
      for (int i = 1; i < it; i++) {
        const double t = cs[i - 1] * h(i - 1, it - 1) + sn[i - 1] * h(i, it - 1);
        h(i, it - 1) = cs[i - 1] * h(i, it - 1) - sn[i - 1] * h(i - 1, it - 1);
        h(i - 1, it - 1) = t;
      }
      if (h(it, it - 1) == 0.0) {
        cs[it - 1] = 1.0; sn[it - 1] = 0.0;
      } else if (std::fabs(h(it, it - 1)) > std::fabs(h(it - 1, it - 1))) {
        const double t = h(it - 1, it - 1) / h(it, it - 1);
        sn[it - 1] = (2.0 * double(h(it, it - 1) > 0) - 1.0) / std::sqrt(1.0 + t * t);
        cs[it - 1] = t * sn[it - 1];
      } else {
        const double t = h(it, it - 1) / h(it - 1, it - 1);
        cs[it - 1] = (2.0 * double(h(it - 1, it - 1) > 0) - 1.0) / std::sqrt(1.0 + t * t);
        sn[it - 1] = t * cs[it - 1];
      }
      h(it - 1, it - 1) = cs[it - 1] * h(it - 1, it - 1) + sn[it - 1] * h(it, it - 1);
      h(it, it - 1) = 0.0;
      beta[it] = -sn[it - 1] * beta[it - 1];
      beta[it - 1] = cs[it - 1] * beta[it - 1];
      const double rel = std::fabs(beta[it]) / nrhs;
      if (rel <= p.gmres_tol || it == maxit) {
        std::vector<double> y(it, 0.0);
        y[it - 1] = beta[it - 1] / h(it - 1, it - 1);
        for (int i = it - 2; i >= 0; i--) {
          y[i] = beta[i];
          for (int j = it - 1; j > i; j--) y[i] -= h(i, j) * y[j];
          y[i] /= h(i, i);
        }
        std::fill(t1.begin(), t1.end(), 0.0);
        for (int i = 0; i < it; i++)
          for (int e = 0; e < n; e++) t1[e] += y[i] * col(i)[e];
        precond(t1.data(), t2.data());
        for (int e = 0; e < n; e++) soln[e] += t2[e];
        unpack(soln);
        return it;
      }
    }
    return maxit;
  }
};

extern "C" {

orc_cilia *orc_cilia_create(const orc_cilia_params *p, int nthreads) {
  if (!p || p->n_swimmers < 1 || p->nseg < 2) return nullptr;
  // config.hpp:276-289: the wall build has one swimmer, no blobs and
  // prescribed (zero) body velocities
  if (p->wall && (p->n_swimmers != 1 || p->nblob[0] != 0 || !p->fixed_body || p->body_velocity ||
                  (p->body_x && (p->body_x[0] != 0.0 || p->body_x[1] != 0.0 || p->body_x[2] != 0.0))))
    return nullptr;
  std::unique_ptr<orc_cilia> s(new orc_cilia());
  s->p = *p;
  s->nthreads = std::max(1, nthreads);
  s->S = p->nseg;
  s->Nsw = p->n_swimmers;
  s->DL = 2.2 * p->a_seg;
  s->F = fourier_table(p->beat);
  if (p->dynamic_phase) {
    if (!p->gen_force_refs || p->n_gen_force_refs < 2) return nullptr;
    s->gref.assign(p->gen_force_refs, p->gen_force_refs + p->n_gen_force_refs);
  }
  s->sw.resize(s->Nsw);
  size_t bo = 0, fo = 0;
  for (int n = 0; n < s->Nsw; n++) {
    CSw &B = s->sw[n];
    B.nfil = p->nfil[n];
    B.nblob = p->nblob[n];
    B.seg0 = s->Ns; B.blob0 = s->Nb; B.fil0 = s->Nf;
    s->Ns += B.nfil * s->S; s->Nb += B.nblob; s->Nf += B.nfil;
    B.blob_ref.assign(p->blob_refs + 3 * bo, p->blob_refs + 3 * (bo + B.nblob));
    bo += B.nblob;
    if (p->body_x) for (int k = 0; k < 3; k++) B.x[k] = p->body_x[3 * n + k];
    if (p->body_velocity) for (int k = 0; k < 6; k++) B.vp[k] = p->body_velocity[6 * n + k];
    B.KI = Mat(6, 6);
  }
  s->fil.resize(s->Nf);
  for (int n = 0; n < s->Nsw; n++) {
    CSw &B = s->sw[n];
    for (int i = 0; i < B.nfil; i++, fo++) {
      CFil &c = s->fil[fo];
      c.sw = n;
      const double *r = p->fil_refs + 3 * fo;
      for (int k = 0; k < 3; k++) c.ref[k] = r[k];
      // filament.cpp:197-282: frame mapping x to the surface normal and the
      // second frame vector to the polar direction; metachronal phase
      const double rn = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
      // swimmer.cpp:91: above the wall every filament stands along +z
      const double dir[3] = {p->wall ? 0.0 : r[0] / rn, p->wall ? 0.0 : r[1] / rn,
                             p->wall ? 1.0 : r[2] / rn};
      Quat qt;
      qt.s = dir[0]; qt.v[0] = 0.0; qt.v[1] = -dir[2]; qt.v[2] = dir[1];
      qt = qsqrt(qt);
      double bn[3];
      normal(qt, bn);
      double pd[3] = {0.0, 0.0, -1.0};
      const double dp = dir[0] * pd[0] + dir[1] * pd[1] + dir[2] * pd[2];
      for (int k = 0; k < 3; k++) pd[k] -= dp * dir[k];
      const double pn = std::sqrt(pd[0] * pd[0] + pd[1] * pd[1] + pd[2] * pd[2]);
      Quat q2;
      if (pn > 0.0) {
        for (int k = 0; k < 3; k++) pd[k] /= pn;
        q2.s = pd[0] * bn[0] + pd[1] * bn[1] + pd[2] * bn[2];
        q2.v[0] = bn[1] * pd[2] - bn[2] * pd[1];
        q2.v[1] = bn[2] * pd[0] - bn[0] * pd[2];
        q2.v[2] = bn[0] * pd[1] - bn[1] * pd[0];
        q2 = qsqrt(q2);
      }
      c.qm1 = qmul(q2, qt);
      c.q = c.qm1;
      if (p->phase) {
        c.phase = p->phase[fo];
      } else if (p->wall) {
        // filament.cpp:222
        c.phase = 2.0 * kPi * r[0] / (s->DL * (s->S - 1)) + 2.0 * kPi * 116.0 / 300.0;
      } else {
        const double base[2] = {B.x[0] + r[0], B.x[1] + r[1]};
        const double phi = std::atan2(base[1], base[0]);
        c.phase = phi * (2.0 * kPi * p->r_over_l);
        c.phase -= 2.0 * kPi * std::floor(0.5 * c.phase / kPi);
      }
      c.omega0 = p->omega0 ? p->omega0[fo] : 2.0 * kPi;
      c.speed = c.omega0;
      c.K.assign(3 * s->S, 0.0);
    }
  }
  s->xs.assign(3 * size_t(s->Ns), 0.0);
  s->xb.assign(3 * size_t(s->Nb), 0.0);
  s->f_segs.assign(3 * size_t(s->Ns), 0.0);
  s->f_blobs.assign(3 * size_t(s->Nb), 0.0);
  s->v_bodies.assign(6 * size_t(s->Nsw), 0.0);
  s->nsys = 3 * s->Ns + 3 * s->Nb + (p->fixed_body ? 0 : 6 * s->Nsw) +
            (p->dynamic_phase ? s->Nf : 0);
  return s.release();
}

void orc_cilia_destroy(orc_cilia *s) { delete s; }

int orc_cilia_system_size(const orc_cilia *s) { return s->nsys; }

// swimmer.cpp:240-275
void orc_cilia_initial_guess(orc_cilia *s, int nt) {
  (void)nt;
  for (int n = 0; n < s->Nsw; n++) {
    CSw &B = s->sw[n];
    B.qm1 = B.q;
    const M3 R = rot_mat(B.q);
    for (int i = 0; i < B.nfil; i++) {
      const int f = B.fil0 + i;
      double r[3];
      mv3(R, s->fil[f].ref, r);
      const double base[3] = {B.x[0] + r[0], B.x[1] + r[1], B.x[2] + r[2]};
      s->filament_guess(f, base, B.u);
    }
    s->make_precon_mat(n);
  }
}

// mobility_solver.cpp:1358-1366 (with read_positions_and_forces' blob
// positions, :132-150)
int orc_cilia_compute_velocities(orc_cilia *s, int nt, int *gmres_iters) {
  s->blob_positions();
  s->assemble_rhs(nt);
  const int it = s->solve();
  s->last_iters = it;
  if (gmres_iters) *gmres_iters = it;
  return 0;
}

// cilia_sim_main.cpp:218-232
void orc_cilia_integrate_bodies(orc_cilia *s) {
  for (int n = 0; n < s->Nsw; n++) {
    CSw &B = s->sw[n];
    const double *V = &s->v_bodies[6 * n];
    for (int k = 0; k < 3; k++) B.x[k] += s->p.dt * V[k];
    for (int k = 0; k < 3; k++) B.u[k] = s->p.dt * V[3 + k];
    B.q = qmul(lie_exp(B.u), B.qm1);
  }
}

// filament.cpp:1020-1031
void orc_cilia_end_of_step(orc_cilia *s, int nt) {
  (void)nt;
  for (CFil &c : s->fil) {
    c.phase += c.speed * s->p.dt;
    c.qm1 = c.q;
  }
}

int orc_cilia_step(orc_cilia *s, int nt, int *gmres_iters) {
  orc_cilia_initial_guess(s, nt);
  const int rc = orc_cilia_compute_velocities(s, nt, gmres_iters);
  if (rc) return rc;
  orc_cilia_integrate_bodies(s);
  orc_cilia_end_of_step(s, nt);
  return 0;
}

void orc_cilia_rhs(orc_cilia *s, double *out) {
  s->blob_positions();
  s->assemble_rhs(0);
  std::copy(s->rhs.begin(), s->rhs.end(), out);
}

void orc_cilia_apply_operator(orc_cilia *s, const double *in, double *out) {
  s->blob_positions();
  s->sys_mult(in, out);
}

void orc_cilia_apply_preconditioner(orc_cilia *s, const double *in, double *out) {
  s->precond(in, out);
}

void orc_cilia_precon_matrices(const orc_cilia *s, double *out36) {
  for (int n = 0; n < s->Nsw; n++)
    std::copy(s->sw[n].KI.d.begin(), s->sw[n].KI.d.end(), out36 + 36 * n);
}

void orc_cilia_get_state(const orc_cilia *s, double *seg_x, double *phase, double *speed,
                         double *body, double *vel_dir, double *fil_q) {
  if (seg_x) std::copy(s->xs.begin(), s->xs.end(), seg_x);
  for (int f = 0; f < s->Nf; f++) {
    const CFil &c = s->fil[f];
    if (phase) phase[f] = c.phase;
    if (speed) speed[f] = c.speed;
    if (vel_dir) std::copy(c.K.begin(), c.K.end(), vel_dir + 3 * size_t(f) * s->S);
    if (fil_q) {
      fil_q[4 * f] = c.q.s;
      for (int k = 0; k < 3; k++) fil_q[4 * f + 1 + k] = c.q.v[k];
    }
  }
  if (body)
    for (int n = 0; n < s->Nsw; n++) {
      const CSw &B = s->sw[n];
      for (int k = 0; k < 3; k++) body[7 * n + k] = B.x[k];
      body[7 * n + 3] = B.q.s;
      for (int k = 0; k < 3; k++) body[7 * n + 4 + k] = B.q.v[k];
    }
}

void orc_cilia_get_solution(const orc_cilia *s, double *f_segs, double *f_blobs,
                            double *v_bodies) {
  if (f_segs) std::copy(s->f_segs.begin(), s->f_segs.end(), f_segs);
  if (f_blobs) std::copy(s->f_blobs.begin(), s->f_blobs.end(), f_blobs);
  if (v_bodies) std::copy(s->v_bodies.begin(), s->v_bodies.end(), v_bodies);
}

void orc_cilia_generalised_forces(const orc_cilia *s, double *out) {
  for (int f = 0; f < s->Nf; f++) {
    double g = 0.0;
    for (int k = 0; k < 3 * s->S; k++) g += s->fil[f].K[k] * s->f_segs[3 * size_t(f) * s->S + k];
    out[f] = g;
  }
}

void orc_cilia_segment_velocities(orc_cilia *s, double *out) {
  std::vector<double> vb(3 * size_t(s->Nb));
  s->blob_positions();
  s->mobility(s->f_segs.data(), s->f_blobs.data(), out, vb.data());
}

double orc_cilia_interp_gen_force(const orc_cilia *s, double phase, double omega0) {
  return s->interp(phase, omega0);
}

}  // extern "C"
