/* gmi_cilia.h -- prescribed-cilia time step on the device (the reference's
 * CILIA_TYPE 3 build with the GMRES solver).
 *
 * Each filament follows a prescribed sequence of shapes indexed by its
 * phase; every step solves one saddle-point system for the segment forces,
 * the blob forces, the body velocities (free bodies) and the phase speeds
 * (dynamic phases) with right-preconditioned GMRES whose operator contains
 * all four translation-translation mobility blocks
 * (mobility_solver.cpp:733-904), then moves the bodies and phases explicitly
 * (cilia_sim_main.cpp:218-232, filament.cpp:1020-1031).
 *
 * Conventions as in gmi.h.  A context carries either this or a gmi_sim,
 * never both; the context's pair launches become translation-only (the
 * reference compiles the rotation terms out, cuda_functions.cu:345,646,808).
 * Vectors of the system in the reference layout (mobility_solver.cpp:
 * 62-110): segment forces [3 Ns] | blob forces [3 Nb] | body velocities
 * [6 Nsw] (free bodies) | phase speeds [Nf] (dynamic phases).
 *
 * Carpet above a no-slip wall (the reference's INFINITE_PLANE_WALL build,
 * config.hpp:276-289): call gmi_set_wall(ctx, 1) on a context with one group
 * of segments and no blobs before gmi_cilia_create.  The surface is at rest
 * (fixed_body = 1, body_velocity = NULL, body_x NULL or the origin),
 * fil_refs are the base positions above the plane (z > 0), every beat plane
 * holds +z, the default
 * phase is 2 pi x / (DL (NSEG - 1)) + 2 pi 116/300 (filament.cpp:222), the
 * mobility is the translation part of the Swan-Brady wall kernel
 * (cuda_functions.cu:8-242 with PRESCRIBED_CILIA), FP64 only. */
#ifndef GMI_CILIA_H
#define GMI_CILIA_H

#include "gmi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Beat shapes (config.hpp:21-37 SHAPE_SEQUENCE) */
#define GMI_BEAT_BUILD 0          /* 'build-a-beat' (filament.cpp:717-815)            */
#define GMI_BEAT_FULFORD_BLAKE 1  /* fitted airway beat (filament.cpp:471-520)        */
#define GMI_BEAT_CORAL_LARVAE 2   /* fitted coral-larva beat (filament.cpp:522-549)   */

/* Replaces the compile-time options of config.hpp:21-76 and the state the
 * reference builds in swimmer::initial_setup / filament::initial_setup
 * (swimmer.cpp:160-237, filament.cpp:197-290).  Arrays need not outlive the
 * call. */
typedef struct gmi_cilia_params {
  const int *nfil;              /* NFIL per swimmer; NFIL*NSEG must equal the
                                   context's segments of that swimmer         */
  const double *blob_refs;      /* body-frame blob positions, all swimmers     */
  const double *fil_refs;       /* body-frame filament base positions          */
  const double *body_x;         /* initial body centres [3 Nsw], NULL: origin  */
  const double *phase;          /* initial phases [Nf]; NULL: the reference's
                                   metachronal start, phase = 2 pi R_OVER_L
                                   atan2(y, x) of the base (filament.cpp:226;
                                   above a wall: filament.cpp:222)           */
  const double *omega0;         /* natural frequencies [Nf]; NULL: 2 pi        */
  double r_over_l;              /* R_OVER_L (default phases only)              */
  double dt;                    /* DT                                          */
  int beat;                     /* GMI_BEAT_*                                  */
  double amplitude;             /* SCALED_BEAT_AMPLITUDE (build-a-beat)        */
  double recovery_window;       /* RECOVERY_STROKE_WINDOW_LENGTH               */
  double effective_stroke;      /* EFFECTIVE_STROKE_LENGTH                     */
  double zero_velocity_avoidance; /* ZERO_VELOCITY_AVOIDANCE_LENGTH            */
  int dynamic_phase;            /* DYNAMIC_PHASE_EVOLUTION                     */
  int fixed_body;               /* PRESCRIBED_BODY_VELOCITIES                  */
  const double *body_velocity;  /* prescribed (V, W) [6 Nsw] for fixed bodies;
                                   NULL: zero (rigid_body.cpp:436-450)         */
  const double *gen_force_refs; /* reference generalised forces sampled evenly
                                   over one period (mobility_solver.cpp:36-60);
                                   required with dynamic_phase               */
  int n_gen_force_refs;         /* >= 2                                        */
  double gmres_tol;             /* LINEAR_SYSTEM_TOL                           */
  int gmres_max_iter;           /* MAX_LINEAR_SYSTEM_ITER, at most 1023        */
} gmi_cilia_params;

/* Swimmer, filament and mobility-solver set-up of the prescribed build
 * (swimmer.cpp:160-237, filament.cpp:197-290, mobility_solver.cpp:22-127). */
int gmi_cilia_create(gmi_ctx *ctx, const gmi_cilia_params *params);

/* One iteration of the time loop (cilia_sim_main.cpp:205-232, 320-330):
 * initial guess, compute_velocities, explicit body update, end_of_step. */
int gmi_cilia_step(gmi_ctx *ctx, int nt, int *gmres_iters);
/* swimmer::initial_guess (swimmer.cpp:240-275): shapes, velocity directions,
 * preconditioner matrices */
int gmi_cilia_initial_guess(gmi_ctx *ctx, int nt);
/* mobility_solver::compute_velocities (mobility_solver.cpp:1358-1366) */
int gmi_cilia_compute_velocities(gmi_ctx *ctx, int nt, int *gmres_iters);
/* explicit body update (cilia_sim_main.cpp:218-232) */
int gmi_cilia_integrate_bodies(gmi_ctx *ctx);
/* swimmer::end_of_step -> filament::end_of_step (filament.cpp:1020-1031) */
int gmi_cilia_end_of_step(gmi_ctx *ctx, int nt);

/* State: segment positions [3 Ns], phases and phase speeds [Nf], bodies
 * [7 Nsw] = (x, q), velocity directions [3 Ns] (filament::vel_dir); any may
 * be NULL. */
int gmi_cilia_get_state(gmi_ctx *ctx, double *seg_x, double *phase, double *phase_speed,
                        double *body, double *vel_dir);
/* The last solution (f_segs_host / f_blobs_host / v_bodies of the
 * reference): [3 Ns], [3 Nb], [6 Nsw]; any may be NULL. */
int gmi_cilia_get_solution(gmi_ctx *ctx, double *f_segs, double *f_blobs, double *v_bodies);
/* vel_dir . f per filament (cilia_sim_main.cpp:234-246) [Nf] */
int gmi_cilia_generalised_forces(gmi_ctx *ctx, double *out);
/* Segment velocities M_ss f_s + M_sb f_b of the last solution [3 Ns]
 * (write_data's system_matrix_mult, mobility_solver.cpp:1777-1785). */
int gmi_cilia_segment_velocities(gmi_ctx *ctx, double *v_segs);

/* Building blocks on the current geometry, reference layout [system size]:
 * assemble_rhs (mobility_solver.cpp:196-359), system_matrix_mult (:733-904),
 * apply_preconditioner (:427-650). */
int gmi_cilia_system_size(gmi_ctx *ctx, int *n);
int gmi_cilia_rhs(gmi_ctx *ctx, double *out);
int gmi_cilia_apply_operator(gmi_ctx *ctx, const double *in, double *out);
int gmi_cilia_apply_preconditioner(gmi_ctx *ctx, const double *in, double *out);

/* GMI_FP64, or GMI_FP32_KAHAN: relaxed-precision GMRES (FP64 operator until
 * the residual is small, FP32 pairs after, FP64 iterative refinement). */
int gmi_cilia_set_precision(gmi_ctx *ctx, int precision);
/* HIP-graph capture of the Arnoldi chunks (default on, chunk 6). */
int gmi_cilia_set_graphs(gmi_ctx *ctx, int use_graphs, int chunk);
/* Gram-Schmidt passes per Arnoldi step: 1 (default, one pass as the
 * reference's modified Gram-Schmidt, mobility_solver.cpp:1175-1183) or 2
 * (reorthogonalised).  Captured graphs are dropped. */
int gmi_cilia_set_gram_schmidt(gmi_ctx *ctx, int passes);

#ifdef __cplusplus
}
#endif
#endif
