/* oracle_sim.h -- CPU FP64 restatement of the reference implicit time step.
 * TEST INFRASTRUCTURE ONLY (see oracle.h). */
#ifndef GMI_ORACLE_SIM_H
#define GMI_ORACLE_SIM_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_sim_params {
  int n_swimmers;
  const int *nfil;          /* filaments per swimmer                        */
  const int *nblob;         /* blobs per swimmer                            */
  int nseg;                 /* NSEG                                         */
  const double *blob_refs;  /* body-frame blob positions, all swimmers      */
  const double *fil_refs;   /* body-frame filament base displacements       */
  const double *body_x;     /* initial body centres [3 n_swimmers] or NULL  */
  const double *fil_phase;  /* driving phase per filament or NULL           */
  double mu, a_seg, a_blob; /* MU, RSEG, RBLOB                              */
  double kb, kt;            /* KB, KT                                       */
  double dt;                /* DT                                           */
  double tol;               /* Broyden TOL                                  */
  double jacobian_confidence;
  double gmres_tol;
  int max_broyden_iter;
  int gmres_max_iter;
  int num_euler_steps;
  double barrier_K;         /* REPULSIVE_FORCE_FACTOR*END_FORCE_MAGNITUDE   */
  double curvature_amp;     /* preferred-curvature amplitude                */
  double curvature_period;
  double sphere_radius;     /* body radius in the Schur RPY correction      */
  const double *sphere_radii; /* per swimmer, or NULL: sphere_radius for all */
  double end_force;         /* END_FORCE_MAGNITUDE of the follower force at
                               each filament tip (CILIA_TYPE 0); 0: none   */
  const double *perturbation; /* per filament 6 values (perturbation1,
                               perturbation2 of filament.cpp:336-360) added
                               to the forces on segments 4 and 12 at nt = 0;
                               NULL: none (needs NSEG >= 13)               */
  int interact;             /* 1: every particle couples to every other in
                               the pair sums and the barrier (the reference's
                               NSWIM > 1 run, cuda_functions.cu:278,457);
                               0: independent swimmers                     */
} orc_sim_params;

typedef struct orc_sim orc_sim;
orc_sim *orc_sim_create(const orc_sim_params *p, int nthreads);
void orc_sim_destroy(orc_sim *s);
int orc_sim_make_body_reference_matrices(orc_sim *s);
int orc_sim_step(orc_sim *s, int nt, int *broyden_iters, int *gmres_iters);
int orc_sim_num_segments(const orc_sim *s);
int orc_sim_num_blobs(const orc_sim *s);
int orc_sim_nbroy(const orc_sim *s);
void orc_sim_get_state(const orc_sim *s, double *seg_x, double *seg_q, double *seg_u,
                       double *fil_lam, double *body_xqu);
void orc_sim_get_mobility(const orc_sim *s, double *v_segs, double *v_bodies, double *f_blobs,
                          double *f_segs, double *body_ref66);
void orc_sim_get_broyden(const orc_sim *s, double *err);
int orc_sim_backup_size(const orc_sim *s);
void orc_sim_get_backup(const orc_sim *s, double *out);
void orc_sim_set_backup(orc_sim *s, const double *in);
int orc_sim_test_compute_velocities(orc_sim *s);
void orc_sim_test_forces(orc_sim *s);
int orc_sim_test_prepare_jacobian(orc_sim *s, int nt);
int orc_sim_test_filament_jacobian(orc_sim *s, int fil, int nt, double *J, double *Jinv);
void orc_sim_test_jacobian_inv_mult(orc_sim *s, const double *in, double *out);
void orc_sim_test_set_driving(orc_sim *s, int nt);
int orc_sim_test_errors(orc_sim *s, int nt, double *out);
void orc_lie_exp(const double *u, double *q);
void orc_lie_midpoint(const double *q1, const double *q2, double *q);
void orc_lie_rotmat(const double *q, double *R);
void orc_lie_dexp(const double *u, const double *v, double *out);
void orc_lie_dexpinv(const double *u, const double *v, double *out);
void orc_lie_dexpinv_transpose(const double *u, const double *v, double *out);
int orc_invert(double *A, int n);
void orc_sim_test_initial_guess(orc_sim *s, int nt);
/* Internal forces and torques of one filament in a given state
 * (filament.cpp:1112-1220): q[4S] (s, v0, v1, v2), u0[3] (segment 0's Lie
 * element, for the clamp torque), lam[3(S-1)], teth[3], clamp[3], strain
 * twist st[3]; end_force and pert (6 values or NULL, applied as at nt = 0).
 * Writes f[6S] (force, torque per segment). */
void orc_test_filament_forces(int S, double DL, double kb, double kt, const double *q,
                              const double *u0, const double *lam, const double *teth,
                              const double *clamp, const double *st, double end_force,
                              const double *pert, double *f);
/* Sets the kinematic state the BDF2 residual reads: per segment 15 values
 * (x, xm1, xm2, u, um1), v_segs[6Ns], per swimmer 15 values (x, xm1, xm2, u,
 * um1) and v_bodies[6 per swimmer]. */
void orc_sim_test_set_kinematics(orc_sim *s, const double *seg, const double *v_segs,
                                 const double *body, const double *v_bodies);
/* Broyden pieces: begin = the step prologue up to the approximate Jacobian
 * (returns 1 if the residual is above tolerance); iterate = find_update,
 * update, forces, velocities, residual, end_of_iter (forced to store the
 * rank-one pair) with dx = the update applied, e_old/e_new the residuals
 * before and after (returns 1 if above tolerance); columns copies C, D
 * (nbroy x max_broyden_iter, row-major) and returns the iteration count. */
int orc_sim_test_broyden_begin(orc_sim *s, int nt);
int orc_sim_test_broyden_iterate(orc_sim *s, int nt, double *dx, double *e_old, double *e_new);
int orc_sim_test_broyden_columns(const orc_sim *s, double *C, double *D);

#ifdef __cplusplus
}
#endif
#endif
