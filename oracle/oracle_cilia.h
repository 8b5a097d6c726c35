/* oracle_cilia.h -- CPU FP64 restatement of the reference's prescribed-cilia
 * time step (CILIA_TYPE 3, GMRES solver).
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Vectors of the saddle-point system use the reference layout
 * (mobility_solver.cpp:62-110): segment forces 3 Ns | blob forces 3 Nb |
 * body velocities 6 Nsw (free bodies only) | phase speeds Nf (dynamic phases
 * only), every block swimmer-major. */
#ifndef GMI_ORACLE_CILIA_H
#define GMI_ORACLE_CILIA_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_cilia_params {
  int n_swimmers;
  const int *nfil;            /* NFIL per swimmer                                */
  const int *nblob;           /* NBLOB per swimmer                               */
  int nseg;                   /* NSEG                                            */
  int interact;               /* 1: all swimmers couple hydrodynamically         */
  const double *blob_refs;    /* body-frame blob positions, all swimmers         */
  const double *fil_refs;     /* body-frame filament base positions              */
  const double *body_x;       /* initial body centres [3 Nsw] or NULL            */
  const double *phase;        /* initial phases [Nf] or NULL (metachronal start) */
  const double *omega0;       /* natural frequencies [Nf] or NULL (2 pi)         */
  double r_over_l;            /* R_OVER_L (default phases)                       */
  double mu, a_seg, a_blob, dt;
  int beat;                   /* 0 build-a-beat, 1 Fulford-Blake, 2 coral larvae */
  double amplitude, recovery_window, effective_stroke, zero_velocity_avoidance;
  int dynamic_phase;          /* DYNAMIC_PHASE_EVOLUTION                         */
  int fixed_body;             /* PRESCRIBED_BODY_VELOCITIES                      */
  const double *body_velocity;/* prescribed (V, W) [6 Nsw] or NULL (zero)        */
  const double *gen_force_refs; /* reference generalised forces over a period    */
  int n_gen_force_refs;
  double gmres_tol;           /* LINEAR_SYSTEM_TOL                               */
  int gmres_max_iter;         /* MAX_LINEAR_SYSTEM_ITER                          */
  int wall;                   /* 1: carpet above a no-slip wall at z = 0 (the
                                 INFINITE_PLANE_WALL build: one swimmer, no
                                 blobs, fixed body at rest, beat frames about
                                 +z, wall phases, wall mobility)               */
} orc_cilia_params;

typedef struct orc_cilia orc_cilia;
orc_cilia *orc_cilia_create(const orc_cilia_params *p, int nthreads);
void orc_cilia_destroy(orc_cilia *s);
int orc_cilia_system_size(const orc_cilia *s);
/* one loop iteration of cilia_sim_main.cpp:205-232 plus end_of_step */
int orc_cilia_step(orc_cilia *s, int nt, int *gmres_iters);
void orc_cilia_initial_guess(orc_cilia *s, int nt);
int orc_cilia_compute_velocities(orc_cilia *s, int nt, int *gmres_iters);
void orc_cilia_integrate_bodies(orc_cilia *s);
void orc_cilia_end_of_step(orc_cilia *s, int nt);
/* building blocks on the current geometry */
void orc_cilia_rhs(orc_cilia *s, double *out);
void orc_cilia_apply_operator(orc_cilia *s, const double *in, double *out);
void orc_cilia_apply_preconditioner(orc_cilia *s, const double *in, double *out);
void orc_cilia_precon_matrices(const orc_cilia *s, double *out36);
/* state: seg_x [3 Ns], phase/speed [Nf], body [7 Nsw] = (x, q), vel_dir [3 Ns],
 * fil_q [4 Nf] (filament frames) -- any may be NULL */
void orc_cilia_get_state(const orc_cilia *s, double *seg_x, double *phase, double *speed,
                         double *body, double *vel_dir, double *fil_q);
/* the last solution: f_segs [3 Ns], f_blobs [3 Nb], v_bodies [6 Nsw] */
void orc_cilia_get_solution(const orc_cilia *s, double *f_segs, double *f_blobs,
                            double *v_bodies);
/* K.f per filament (cilia_sim_main.cpp:234-246) */
void orc_cilia_generalised_forces(const orc_cilia *s, double *out);
/* segment velocities M_ss f_s + M_sb f_b of the last solution [3 Ns] */
void orc_cilia_segment_velocities(orc_cilia *s, double *out);
/* the interpolated generalised force at `phase` (mobility_solver.cpp:223-306),
 * already scaled by omega0 / 2 pi */
double orc_cilia_interp_gen_force(const orc_cilia *s, double phase, double omega0);

#ifdef __cplusplus
}
#endif
#endif
