/* gmi.h -- C ABI of the MI355X (gfx950) RPY grand-mobility library.
 *
 * Drop-in boundary for the reference's mobility backend.  The reference
 * exposes the abstract C++ class `mobility_solver` (mobility_solver.hpp:18-102)
 * whose pure virtuals (mobility_solver.hpp:25-44) are implemented by
 * `rpy_mobility_solver` (rpy_mobility_solver.cu:14-311).  Each entry point
 * below replaces one of those members; the comment on each names it.
 *
 * Conventions
 *  - Plain C types only; every call returns GMI_OK (0) or a negative
 *    GMI_ERR_* code, and gmi_last_error() gives the message.  No exceptions
 *    cross this boundary.
 *  - Host arrays are caller-owned FP64 AoS in the reference layout
 *    (mobility_solver.hpp:54-60): x_segs[3 Ns], f_segs[6 Ns] = (f, tau),
 *    v_segs[6 Ns] = (v, omega), x_blobs/f_blobs/v_blobs[3 Nb]; index order
 *    swimmer-major, then filament, then segment (blobs: swimmer, blob).
 *  - Device buffers are owned by the context.  Work is queued on the
 *    context's stream; copies to host complete at gmi_wait_for_device(),
 *    which is the only synchronisation point (as wait_for_device(),
 *    rpy_mobility_solver.cu:230-239).
 *  - One context per device per host thread.
 */
#ifndef GMI_H
#define GMI_H

#ifdef __cplusplus
extern "C" {
#endif

#define GMI_OK 0
#define GMI_ERR_ARG (-1)        /* bad argument / size                        */
#define GMI_ERR_HIP (-2)        /* HIP runtime error                          */
#define GMI_ERR_STATE (-3)      /* call out of order                          */
#define GMI_ERR_NOT_FINITE (-4) /* NaN/Inf in a result that must be finite    */
#define GMI_ERR_SINGULAR (-5)   /* singular matrix (matrix.cpp:549-553)       */

/* Arithmetic of the pair kernels. */
#define GMI_FP64 0       /* FP64 throughout (parity path)                    */
#define GMI_FP32_KAHAN 1 /* FP32 pairs, Kahan cross-tile sums, FP64 reduce    */

typedef struct gmi_ctx gmi_ctx;

/* Sizes and physics of one context (replaces the compile-time NSWIM, NFIL,
 * NSEG, NBLOB, MU, RSEG, RBLOB of config.hpp:95-103). */
typedef struct gmi_params {
  int n_swimmers;               /* NSWIM                                      */
  const int *nseg_per_swimmer;  /* NFIL*NSEG of each swimmer [n_swimmers]     */
  const int *nblob_per_swimmer; /* NBLOB of each swimmer [n_swimmers]         */
  int nseg_per_filament;        /* NSEG (barrier exclusion, filament layout)  */
  int swimmers_interact;        /* 1: every particle couples to every other
                                   (the reference's NSWIM semantics,
                                   cuda_functions.cu:278,457); 0: swimmers
                                   are hydrodynamically independent          */
  double mu;                    /* MU                                         */
  double a_seg;                 /* RSEG                                       */
  double a_blob;                /* RBLOB                                      */
  int precision;                /* GMI_FP64 or GMI_FP32_KAHAN                 */
  int device;                   /* HIP device ordinal                         */
} gmi_params;

/* Number of visible HIP devices (cudaGetDeviceCount in
 * allocate_host_memory, rpy_mobility_solver.cu:58). */
int gmi_device_count(int *count);
/* allocate_host_memory + allocate_device_memory (rpy_mobility_solver.cu:54-115) */
int gmi_create(gmi_ctx **ctx, const gmi_params *params);
/* free_host_memory + free_device_memory (rpy_mobility_solver.cu:14-52) */
void gmi_destroy(gmi_ctx *ctx);
const char *gmi_last_error(const gmi_ctx *ctx);
int gmi_set_precision(gmi_ctx *ctx, int precision);
int gmi_num_segments(const gmi_ctx *ctx);
int gmi_num_blobs(const gmi_ctx *ctx);

/* ---- host <-> device copies (rpy_mobility_solver.cu:119-205) ------------ */
/* copy_segment_positions_to_device (rpy_mobility_solver.cu:119) */
int gmi_copy_segment_positions_to_device(gmi_ctx *ctx, const double *x_segs);
/* copy_segment_forces_to_device (rpy_mobility_solver.cu:130) */
int gmi_copy_segment_forces_to_device(gmi_ctx *ctx, const double *f_segs);
/* copy_blob_positions_to_device (rpy_mobility_solver.cu:141) */
int gmi_copy_blob_positions_to_device(gmi_ctx *ctx, const double *x_blobs);
/* copy_blob_forces_to_device (rpy_mobility_solver.cu:152) */
int gmi_copy_blob_forces_to_device(gmi_ctx *ctx, const double *f_blobs);
/* copy_interparticle_blob_forces_to_host (rpy_mobility_solver.cu:163) */
int gmi_copy_interparticle_blob_forces_to_host(gmi_ctx *ctx, double *f_blobs_rep);
/* copy_blob_velocities_to_host (rpy_mobility_solver.cu:177) */
int gmi_copy_blob_velocities_to_host(gmi_ctx *ctx, double *v_blobs);
/* copy_segment_velocities_to_host (rpy_mobility_solver.cu:191) */
int gmi_copy_segment_velocities_to_host(gmi_ctx *ctx, double *v_segs);

/* ---- device work (rpy_mobility_solver.cu:207-311) ----------------------- */
/* apply_interparticle_forces (rpy_mobility_solver.cu:207-228): steric
 * barrier; adds into the device segment forces and sets the device blob
 * repulsion.  K = REPULSIVE_FORCE_FACTOR*END_FORCE_MAGNITUDE (config.hpp:239-247). */
int gmi_apply_interparticle_forces(gmi_ctx *ctx, double K);
/* wait_for_device (rpy_mobility_solver.cu:230-239) */
int gmi_wait_for_device(gmi_ctx *ctx);
/* evaluate_segment_segment_mobility -> Mss_mult: V_s = M_ss F_s (rpy_mobility_solver.cu:241-257) */
int gmi_evaluate_segment_segment_mobility(gmi_ctx *ctx);
/* evaluate_segment_blob_mobility -> Msb_mult: V_s += M_sb F_b (rpy_mobility_solver.cu:259-275) */
int gmi_evaluate_segment_blob_mobility(gmi_ctx *ctx);
/* evaluate_blob_blob_mobility -> Mbb_mult: V_b = M_bb F_b (rpy_mobility_solver.cu:277-293) */
int gmi_evaluate_blob_blob_mobility(gmi_ctx *ctx);
/* evaluate_blob_segment_mobility -> Mbs_mult: V_b = M_bs F_s (rpy_mobility_solver.cu:295-311) */
int gmi_evaluate_blob_segment_mobility(gmi_ctx *ctx);
/* All four blocks in one launch: V_s = M_ss F_s + M_sb F_b,
 * V_b = M_bs F_s + M_bb F_b.  No single reference counterpart; equals the
 * sum of the four evaluate_* calls above. */
int gmi_evaluate_grand_mobility(gmi_ctx *ctx);

/* ---- timing (no reference counterpart; the reference has no timers) ----- */
/* Records a HIP event on the context's stream. */
int gmi_timer_start(gmi_ctx *ctx);
/* Records the stop event, waits for it, returns elapsed milliseconds. */
int gmi_timer_stop(gmi_ctx *ctx, double *ms);

/* Starts recording HIP events around every pair-kernel launch (the
 * reduction of split partials excluded). */
int gmi_profile_begin(gmi_ctx *ctx);
/* Waits for the stream, returns the summed pair-kernel time in ms and the
 * number of launches since gmi_profile_begin(), and stops recording. */
int gmi_profile_end(gmi_ctx *ctx, double *pair_ms, int *launches);
/* While profiling, every eagerly launched Arnoldi step (graphs off) also
 * records HIP events around its Gram-Schmidt kernels (both passes: column
 * dots, subtraction, norm).  Waits for the stream and returns, for the steps
 * recorded since gmi_profile_begin() or the previous call, the summed time
 * in ms, the algorithmic bytes 2 x 8 (k + 2) n of each step (k previous
 * columns of length n: per pass the columns and w read, w written), and the
 * step count; then clears them.  Steps launched past convergence are not
 * counted.  The reference times nothing; its Gram-Schmidt is
 * mobility_solver.cpp:1175-1183. */
int gmi_profile_krylov(gmi_ctx *ctx, double *gs_ms, double *gs_bytes, int *steps);

/* Segment-segment mobility above a no-slip plane wall at z = 0 (the
 * reference's INFINITE_PLANE_WALL build: Swan & Brady, cuda_functions.cu:
 * 8-242).  wall != 0 switches gmi_evaluate_segment_segment_mobility (and the
 * grand apply, which then has no other blocks) to it; FP64 whatever the
 * context precision.  The reference's wall build has no bodies: contexts
 * with blobs are refused.  Set it before gmi_cilia_create to step a carpet
 * of prescribed cilia above the wall (gmi_cilia.h); the setting is fixed
 * while a time stepper exists, and gmi_sim_create refuses a wall.  Segments
 * must lie above the wall (z > 0). */
int gmi_set_wall(gmi_ctx *ctx, int wall);

/* ---- launch shaping (no reference counterpart; the reference launches
 * ceil(N/64) blocks of 64, rpy_mobility_solver.cu:249-305) --------------- */
/* Pair launches aim at `target_blocks` workgroups by splitting each target
 * tile's sources into chunks of at least `min_chunk` sources. */
int gmi_set_launch_shape(gmi_ctx *ctx, int target_blocks, int min_chunk);
/* Times the grand apply on the context's current device data for a grid of
 * launch shapes (reps launches each), applies the fastest and reports it.
 * The choice is cached per (sizes, precision) for the process. */
int gmi_autotune(gmi_ctx *ctx, int reps, int *best_target_blocks, int *best_min_chunk);
/* The same for the launches of one block subset: mask is an OR of GMI_BLOCK_*
 * (GMI_BLOCK_BB alone is the operator of the rigid-body GMRES). */
#define GMI_BLOCK_SS 1 /* segment <- segment (Mss) */
#define GMI_BLOCK_SB 2 /* segment <- blob (Msb)    */
#define GMI_BLOCK_BS 4 /* blob <- segment (Mbs)    */
#define GMI_BLOCK_BB 8 /* blob <- blob (Mbb)       */
int gmi_autotune_blocks(gmi_ctx *ctx, int mask, int reps, int *best_target_blocks,
                        int *best_min_chunk);

/* ---- grand apply that allocates nothing ------------------------------- */
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  It is gmi_evaluate_grand_mobility (same kernels, same
 * source splits as a context with the default launch shape) on arrays the
 * caller owns, with no context:
 *   - every pointer is DEVICE memory of the current device, SoA: component c
 *     of particle i at p[c*ld + i] (x_*: 3 rows, f_segs: 6 rows (f, tau),
 *     f_blobs: 3 rows, v_segs: 6 rows (v, omega), v_blobs: 3 rows), ld >= n;
 *   - v_segs = M_ss f_segs + M_sb f_blobs, v_blobs = M_bs f_segs + M_bb f_blobs
 *     over all particles (one interacting group), precision GMI_FP64 or
 *     GMI_FP32_KAHAN;
 *   - work is queued on `stream` (a hipStream_t; NULL: the null stream) and
 *     is complete when that stream is;
 *   - no hipMalloc/hipFree and no allocating library call: the tile table,
 *     work list and split partials live in a fixed device array of
 *     GMI_NOALLOC_SCRATCH_BYTES per device, taken on the first call on a
 *     device (make that first call outside graph capture) and held until the
 *     process ends;
 *   - limit: ceil(n_segs/256) + ceil(n_blobs/256) <= GMI_NOALLOC_MAX_TILES
 *     (at least 1,048,064 particles); larger inputs return GMI_ERR_TOO_LARGE
 *     before any device work; devices numbered >= GMI_NOALLOC_MAX_DEVICES
 *     return GMI_ERR_ARG;
 *   - all calls on one device share its array: calls that overlap on one
 *     device (several streams or host threads) race; calls queued on one
 *     stream are ordered. */
#define GMI_ERR_TOO_LARGE (-7)
#define GMI_NOALLOC_MAX_TILES 4096
#define GMI_NOALLOC_MAX_DEVICES 64
#define GMI_NOALLOC_SCRATCH_BYTES 31637504
int gmi_evaluate_grand_mobility_noalloc(
    const double *x_segs, const double *f_segs, int n_segs, int ld_segs,
    const double *x_blobs, const double *f_blobs, int n_blobs, int ld_blobs,
    double *v_segs, double *v_blobs, double mu, double a_seg, double a_blob, int precision,
    void *stream);

#ifdef __cplusplus
}
#endif
#endif /* GMI_H */
