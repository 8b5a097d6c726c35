/* gmi_sim.h -- C ABI of the device-resident implicit time step.
 *
 * Replaces the generic half of `mobility_solver` (mobility_solver.cpp) and
 * the model/solver objects that drive it (swimmer.cpp, filament.cpp,
 * segment.cpp, rigid_body.cpp, broyden_solver.cpp, cilia_sim_main.cpp:205-341).
 * The whole state -- segment positions/orientations, Lagrange multipliers,
 * body state, Broyden C/D matrices, GMRES Krylov basis -- lives on the
 * device; the host only reads convergence flags.  Every entry point below
 * names the reference member it replaces.
 *
 * Swimmers may differ in filament and blob counts; they share NSEG and the
 * radii.  A sim is attached to a gmi_ctx created with the same n_swimmers and
 * per-swimmer sizes (nseg_per_swimmer = nfil*NSEG).  With swimmers_interact
 * = 1 every particle couples to every other, as in the reference's NSWIM > 1
 * runs (cuda_functions.cu:278,457); with 0 the mobility is block-diagonal
 * (independent swimmers stepped as one batch).  The preconditioners and the
 * fitted GMRES guess act per swimmer either way; between coupled swimmers
 * they are no longer exact, so GMRES takes more iterations.
 */
#ifndef GMI_SIM_H
#define GMI_SIM_H

#include "gmi.h"

#ifdef __cplusplus
extern "C" {
#endif

#define GMI_ERR_NOT_CONVERGED (-6) /* Broyden hit MAX_BROYDEN_ITER (cilia_sim_main.cpp:328-335) */

typedef struct gmi_sim_params {
  const int *nfil;            /* NFIL of each swimmer                        */
  const double *blob_refs;    /* body-frame blob positions [3 Nb], swimmer-major
                                 (rigid_body::blob_references)               */
  const double *fil_refs;     /* body-frame filament base displacements
                                 [3 Nfil] (swimmer::filament_references)     */
  const double *body_x;       /* initial body centres [3 n_swimmers] or NULL */
  const double *fil_phase;    /* driving phase per filament or NULL          */
  double kb, kt;              /* KB, KT (config.hpp:105-106)                 */
  double dt;                  /* DT                                          */
  double tol;                 /* Broyden TOL (config.hpp:143)                */
  double jacobian_confidence; /* JACOBIAN_CONFIDENCE_FACTOR (config.hpp:140) */
  double gmres_tol;           /* LINEAR_SYSTEM_TOL (config.hpp:154)          */
  int max_broyden_iter;       /* MAX_BROYDEN_ITER (config.hpp:142)           */
  int gmres_max_iter;         /* MAX_LINEAR_SYSTEM_ITER (config.hpp:153),
                                 at most 1023                                */
  int num_euler_steps;        /* NUM_EULER_STEPS (config.hpp:167)            */
  double barrier_K;           /* REPULSIVE_FORCE_FACTOR*END_FORCE_MAGNITUDE  */
  double curvature_amp;       /* preferred curvature amplitude (strain_twist) */
  double curvature_period;
  double sphere_radius;       /* body radius of the Schur-complement RPY
                                 correction (swimmer.cpp:580)                */
  const double *sphere_radii; /* per-swimmer body radius [n_swimmers], or
                                 NULL for sphere_radius on every swimmer (the
                                 reference has one R_OVER_L, config.hpp:126)  */
  double end_force;           /* END_FORCE_MAGNITUDE (config.hpp:239): the
                                 instability-driven cilia of the default
                                 CILIA_TYPE 0 build -- a follower force
                                 -end_force t at each filament tip, with its
                                 Jacobian and Schur terms (filament.cpp:
                                 1177-1202,1466-1470; swimmer.cpp:497-503,
                                 557-571,630-633); 0: none                   */
  const double *perturbation; /* [6 Nfil] or NULL: per filament the forces
                                 perturbation1, perturbation2 added to
                                 segments 4 and 12 at nt = 0 in the end-force
                                 build (filament.cpp:336-360,1186-1199; the
                                 reference draws them from std::random_device,
                                 here they are an input); needs NSEG >= 13   */
} gmi_sim_params;

/* swimmers[n].initial_setup for all n + mobility.initialise()
 * (swimmer.cpp:13-237, filament.cpp:58-340, rigid_body.cpp:13-245,
 * mobility_solver.cpp:22-130).  Straight filaments along the base normals. */
int gmi_sim_create(gmi_ctx *ctx, const gmi_sim_params *params);
/* make_body_reference_matrices (mobility_solver.cpp:1693-1771), per swimmer. */
int gmi_make_body_reference_matrices(gmi_ctx *ctx);
/* swimmer::initial_guess for every swimmer (swimmer.cpp:240-274). */
int gmi_initial_guess(gmi_ctx *ctx, int nt);
/* swimmer::forces_and_torques for every swimmer (swimmer.cpp:276-325). */
int gmi_forces_and_torques(gmi_ctx *ctx, int nt);
/* mobility_solver::compute_velocities (mobility_solver.cpp:1358-1402):
 * barrier, M_ss/M_bs, rigid-body GMRES, M_sb.  gmres_iters may be NULL. */
int gmi_compute_velocities(gmi_ctx *ctx, int nt, int *gmres_iters);
/* mobility_solver::compute_errors into broyden.error (which = 0) or
 * broyden.new_error (which = 1); *too_large as returned by the reference,
 * *not_finite = 1 if any entry is NaN/Inf (mobility_solver.cpp:1404-1689). */
int gmi_compute_errors(gmi_ctx *ctx, int nt, int which, int *too_large, int *not_finite);
/* swimmer::prepare_jacobian_inv for every swimmer (swimmer.cpp:329-688). */
int gmi_prepare_jacobian_inv(gmi_ctx *ctx, int nt);
/* broyden_solver::find_update (broyden_solver.cpp:26-68). */
int gmi_broyden_find_update(gmi_ctx *ctx);
/* swimmer::update(&broyden.update[...]) for every swimmer (swimmer.cpp:847-871). */
int gmi_apply_update(gmi_ctx *ctx);
/* broyden_solver::end_of_iter (broyden_solver.cpp:70-130). */
int gmi_broyden_end_of_iter(gmi_ctx *ctx, int too_large);
/* One full time step: the body of the nt loop, cilia_sim_main.cpp:205-341.
 * Returns GMI_OK, GMI_ERR_NOT_FINITE (divergence) or GMI_ERR_NOT_CONVERGED. */
int gmi_step(gmi_ctx *ctx, int nt, int *broyden_iters, int *gmres_iters);

/* State access (host arrays).  seg_x[3Ns], seg_q[4Ns] (s, v), seg_u[3Ns];
 * fil_lam[Nfil * (3 NSEG + 3)] = lambda[3(NSEG-1)], tether[3], clamp[3];
 * body_xqu[10 n_swimmers] = x[3], q[4], u[3].  NULL pointers are skipped. */
int gmi_get_state(gmi_ctx *ctx, double *seg_x, double *seg_q, double *seg_u,
                  double *fil_lam, double *body_xqu);
/* v_segs[6Ns] (AoS), v_bodies[6 n_swimmers], f_blobs[3Nb] (GMRES solution,
 * AoS), body_ref66[72 n_swimmers] = N_ref then (K^T K)^-1 per swimmer,
 * row-major.  NULL pointers are skipped. */
int gmi_get_mobility(gmi_ctx *ctx, double *v_segs, double *v_bodies, double *f_blobs,
                     double *body_ref66);
/* broyden.error (which = 0) / new_error (1) / update (2) [NBROY]. */
int gmi_get_broyden(gmi_ctx *ctx, int which, double *out);
/* Solution accuracy of the rigid-body solve: GMI_FP64, or GMI_FP32_KAHAN for
 * FP32 pair kernels inside FP64 iterative refinement (the outer residual
 * uses the FP64 kernel; GMRES tolerance is met in FP64). */
int gmi_sim_set_precision(gmi_ctx *ctx, int precision);
/* GMRES runs its Arnoldi steps in chunks of `chunk` iterations between host
 * convergence checks; with use_graphs each chunk is captured once as a HIP
 * graph and replayed (no reference counterpart).  Default: graphs on, 6.
 * With use_graphs, when GMRES runs in FP64 (GMI_FP64, or under
 * GMI_PRECOND_BODY_INVERSE), each Broyden iteration after the first
 * residual (cilia_sim_main.cpp:266-314) is one captured graph too -- update, forces, velocities, the first GMRES chunk
 * and, gated on the device, the rest of the solve, the residual and the
 * Broyden pair -- and the host reads the flags once per iteration. */
int gmi_sim_set_graphs(gmi_ctx *ctx, int use_graphs, int chunk);
/* Broyden iterations run from captured graphs since the context was created,
 * and the captured graphs the context holds (Arnoldi chunks and Broyden
 * iterations).  No reference counterpart. */
int gmi_sim_graph_stats(gmi_ctx *ctx, int *captured_iterations, int *graphs);
/* GMRES initial guess of the rigid-body solve.  m = 0 (default): the
 * previous solution, as the reference (mobility_solver.cpp:1060-1156).
 * m in [1, 16]: in the body frame, where the system matrix is fixed, the new
 * right-hand side is fitted by least squares onto the swimmer's last m
 * right-hand sides, and the same combination of their solutions is the
 * guess.  GMRES still iterates to the same tolerance.  No reference
 * counterpart: a warm start of this build. */
int gmi_sim_set_gmres_history(gmi_ctx *ctx, int m);

/* Gram-Schmidt passes of each Arnoldi step of the rigid-body GMRES: 1
 * (default) orthogonalises the new vector once against the basis, as the
 * reference's modified Gram-Schmidt does (mobility_solver.cpp:1175-1183); 2
 * repeats the pass (classical Gram-Schmidt with reorthogonalisation).
 * Captured graphs are dropped.  No reference counterpart beyond the choice
 * of algorithm. */
int gmi_sim_set_gram_schmidt(gmi_ctx *ctx, int passes);

/* How the filaments' approximate Jacobians are inverted (the reference's
 * LAPACK getrf/getri, matrix.cpp:538-641).  0 (default): Gauss-Jordan
 * without pivoting, each inverse checked on a fixed vector and redone with
 * partial pivoting if the check fails (NSEG <= 21; longer filaments always
 * pivot).  1: partial pivoting for every filament.  No reference
 * counterpart beyond the choice of algorithm. */
int gmi_sim_set_jacobian_pivoting(gmi_ctx *ctx, int always_pivot);

/* Preconditioner of the rigid-body GMRES. */
#define GMI_PRECOND_SELF_MOBILITY 0 /* the reference's: M ~ I/(6 pi mu a)
                                       (mobility_solver.cpp:651-727)         */
#define GMI_PRECOND_BODY_INVERSE 1  /* the exact block inverse in the body
                                       frame, where the matrix is fixed: M^-1
                                       (3 nb x 3 nb doubles per swimmer, kept
                                       in HBM), M^-1 K and (K^T M^-1 K)^-1,
                                       computed once by block Gauss-Jordan */
/* Selects the preconditioner (set-up cost and memory for BODY_INVERSE:
 * 2 (3 nb)^3 flops and 8 (3 nb)^2 bytes per swimmer during set-up, about half
 * of that resident after it (the packed lower triangle), plus the preconditioned
 * Arnoldi directions, 8 (maxit + 1) rows bytes).  GMRES keeps its operator
 * and tolerance.  BODY_INVERSE applies M^-1 by a symmetric product over its
 * lower triangle and checks convergence after every Arnoldi step (chunk 1;
 * SELF_MOBILITY restores chunk 6; gmi_sim_set_graphs can change it after).
 * No reference counterpart beyond the default. */
int gmi_sim_set_body_preconditioner(gmi_ctx *ctx, int kind);

/* ---- seeding points on the unit sphere (seeding.cu:13-389) ------------- */
#define GMI_SEED_UNIFORM 0    /* UNIFORM_SEEDING: spiral start, then the
                                 repulsion relaxation (seeding.cu:13-130,
                                 232-375; cuda_functions.cu:1159-1250), on
                                 the GPU for any N                          */
#define GMI_SEED_EQUATORIAL 1 /* EQUATORIAL_SEEDING band (seeding.cu:140-175) */
#define GMI_SEED_PLATY 2      /* PLATY_SEEDING band + ring (seeding.cu:177-230) */
/* Fills unit_xyz[3n] on the ctx's device.  r_over_l (R_OVER_L) is used by
 * the band seedings only.  The relaxation stops when no coordinate moves by
 * more than 1e-4 (*iterations reports how many it took) or returns
 * GMI_ERR_NOT_CONVERGED after max_iterations.  The reference caches the
 * result as sphere<N>.seed; generic_mobility_interface_class_amd.geometry
 * reads and writes those files. */
int gmi_seed_sphere(gmi_ctx *ctx, int n, int kind, double r_over_l, int max_iterations,
                    double *unit_xyz, int *iterations);

/* ---- output and checkpoint files in the reference's text formats.  `name`
 * is SIMULATION_NAME (a path prefix); files as in config.hpp:208-215. ---- */
/* <name>.par (cilia_sim_main.cpp:79-116); the format describes identical
 * swimmers, so uneven batches are refused. */
int gmi_sim_write_par(gmi_ctx *ctx, const char *name, int total_time_steps,
                      int plot_frequency_in_steps);
/* <name>_blob_references.dat and <name>_fil_references.dat of swimmer 0
 * (swimmer.cpp:885-900, rigid_body.cpp:383-400, cilia_sim_main.cpp:155). */
int gmi_sim_write_reference_positions(gmi_ctx *ctx, const char *name);
/* Appends one line per file after step nt (cilia_sim_main.cpp:346-366):
 * _seg_states.dat and _body_states.dat start with nt+1 (filament.cpp:
 * 1577-1591, rigid_body.cpp:408-412); _body_vels.dat, _blob_forces.dat,
 * _seg_vels.dat and _seg_forces.dat start with nt (mobility_solver.cpp:
 * 1775-1895).  Values "%.10e " (std::scientific, precision 10). */
int gmi_sim_write_data(gmi_ctx *ctx, int nt, const char *name);
/* Overwrites <name>.backup after step nt (cilia_sim_main.cpp:368-381,
 * swimmer.cpp:914-925, filament.cpp:1593-1650, rigid_body.cpp:414-432).
 * extended != 0 also writes <name>.backup_gmi (no reference counterpart):
 * the same state at full precision plus the GMRES warm start, the previous
 * velocities and the learned warm start's lag features, from which a resumed
 * run repeats the original bit for bit. */
int gmi_sim_write_backup(gmi_ctx *ctx, int nt, const char *name, int extended);
/* Loads <name>.backup into the sim (cilia_sim_main.cpp:29-71, filament.cpp:
 * 84-163, rigid_body.cpp:28-56; q_m1 = q; the GMRES warm start is left as
 * it is, i.e. the last solve's solution, as in a resumed reference run), then
 * <name>.backup_gmi if present (*exact = 1).  *nt_of_backup is the stored
 * step count: continue with gmi_step(ctx, *nt_of_backup, ...) as the
 * reference does for INITIAL_CONDITIONS_TYPE 1 (cilia_sim_main.cpp:193-197).
 * The sizes must match this sim's. */
int gmi_sim_read_backup(gmi_ctx *ctx, const char *name, int *nt_of_backup, int *exact);

/* ---- learned warm start (no reference counterpart: the reference's warm
 * starts are the extrapolations of segment.cpp:53-73, filament.cpp:908-1016,
 * rigid_body.cpp:251-271).  A per-segment MLP (51 -> hidden -> hidden -> 6,
 * tanh, BF16 weights, FP32 accumulation) corrects the extrapolated Broyden
 * unknowns before the first residual of a step: each segment's multiplier
 * triple, and segment 0's clamp multiplier.  Its inputs are 27 state
 * features and the corrections that would have been exact on the previous
 * 4 steps.  Broyden/GMRES still converge to the same tolerances afterwards;
 * an attempt that does not converge in about twice the recent iteration
 * count is redone from the extrapolated guess. */
#define GMI_WARMSTART_NIN 51
/* Creates the network with Glorot-uniform weights from `seed`. */
int gmi_warmstart_init(gmi_ctx *ctx, int hidden, unsigned long long seed);
/* apply: use the trained network in gmi_step (steps nt >= 1);
 * collect_capacity > 0: record (features, converged - guess) rows of the
 * converged steps nt >= 1 with nt % collect_stride == 0 into a dataset of
 * that many rows (0: stop recording). */
int gmi_warmstart_configure(gmi_ctx *ctx, int apply, int collect_capacity, int collect_stride);
int gmi_warmstart_dataset_size(gmi_ctx *ctx, int *rows);
/* Copies the first `rows` recorded rows (raw, not standardised) to the host:
 * X[rows x 51] features, T[rows x 6] converged - guess.  Rows are in
 * recording order: per recorded step, segment-major over all filaments. */
int gmi_warmstart_dataset(gmi_ctx *ctx, float *X, float *T, int rows);
/* Steps that used the network, and those of them whose learned attempt did
 * not converge within about twice the recent Broyden iteration count and
 * were redone from the extrapolated guess (both counted since init). */
int gmi_warmstart_stats(gmi_ctx *ctx, int *applied, int *fallbacks);
/* Adam on the MSE of standardised targets; returns the mean loss of the
 * last 20 iterations. */
int gmi_warmstart_train(gmi_ctx *ctx, int iters, int batch, double lr,
                        unsigned long long seed, double *final_loss);
int gmi_warmstart_param_count(gmi_ctx *ctx, int *n);
/* FP32 master parameters [W1, b1, W2, b2, W3, b3] (row-major, W as
 * inputs x outputs); get and/or set (set re-rounds the BF16 copy). */
int gmi_warmstart_params(gmi_ctx *ctx, float *get, const float *set);
/* Forward pass on B standardised rows X[B x 51] -> Y[B x 6]. */
int gmi_warmstart_forward(gmi_ctx *ctx, const float *X, int B, float *Y);
/* MSE loss of B rows and its gradient w.r.t. the parameters (through the
 * BF16 weights), grad[param_count]. */
int gmi_warmstart_loss_grad(gmi_ctx *ctx, const float *X, const float *T, int B,
                            double *loss, float *grad);

#ifdef __cplusplus
}
#endif
#endif /* GMI_SIM_H */
