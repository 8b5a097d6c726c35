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

#ifdef __cplusplus
}
#endif
#endif
