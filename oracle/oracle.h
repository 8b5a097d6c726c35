/* oracle.h -- CPU FP64 restatement of the reference RPY hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this code.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker / CPU baseline.
 *
 * Layouts follow the reference host buffers (mobility_solver.hpp:54-60):
 * AoS FP64, segments x[3i], f[6i] = (f, tau), v[6i] = (v, omega); blobs
 * x[3m], f[3m], v[3m].  Every function cites the reference file:line it
 * restates.
 */
#ifndef GMI_ORACLE_H
#define GMI_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- pair sums (cuda_functions.cu) ------------------------------------ */
/* Each computes the rows of targets [i0, i1) against every source; V is
 * indexed by absolute target index.  nthreads = OpenMP threads.            */
/* V_s  = M_ss F_s     (cuda_functions.cu:246-425, '=' at :414-419)        */
void orc_Mss(double *V, const double *F, const double *X, int ns, int i0,
             int i1, double a, double mu, int nthreads);
/* V_b  = M_bb F_b     (cuda_functions.cu:429-539, '=' at :531-533)        */
void orc_Mbb(double *V, const double *F, const double *X, int nb, int i0,
             int i1, double a, double mu, int nthreads);
/* V_s += M_sb F_b     (cuda_functions.cu:541-694, '+=' at :683-688)       */
void orc_Msb(double *V, const double *Fb, const double *Xs, int ns,
             const double *Xb, int nb, int i0, int i1, double a_s,
             double a_b, double mu, int nthreads);
/* V_b  = M_bs F_s     (cuda_functions.cu:697-863, '=' at :853-855)        */
void orc_Mbs(double *V, const double *Fs, const double *Xb, int nb,
             const double *Xs, int ns, int i0, int i1, double a_s,
             double a_b, double mu, int nthreads);

/* Steric barrier (cuda_functions.cu:961-1157).  f_segs[6i..6i+2] += F_i,
 * f_blobs_rep[3m..] = F_m.  nseg = segments per filament (exclusion
 * |i-j|<=1 on the same filament), K = REPULSIVE_FORCE_FACTOR*END_FORCE.    */
void orc_barrier(double *f_segs, double *f_blobs_rep, const double *xs, int ns,
                 const double *xb, int nb, int nseg, double a_s, double a_b,
                 double K, int nthreads);

/* Dense 6x6 block (target i, source j) of the equal-radius grand mobility,
 * row-major out[36] mapping (f_j, tau_j) -> (v_i, omega_i).  Used only by
 * the brute-force assembly tests.                                          */
void orc_ss_block(double *out, const double *xi, const double *xj, double a,
                  double mu);

/* Segment-segment mobility above a no-slip wall at z = 0 (Swan & Brady;
 * cuda_functions.cu:8-242), V[6 ns] rows i0..i1. */
void orc_Mss_wall(double *V, const double *F, const double *X, int ns, int i0, int i1,
                  double a, double mu, int nthreads);

/* Unit-sphere seeding (seed.cpp): uniform repulsion (order 1: the
 * reference's GPU summation, 0: its CPU pair loop), returns iterations or -1;
 * and the equatorial / platy bands. */
int orc_seed_uniform(int N, int order, int max_it, double *X);
void orc_seed_band(int N, int platy, double r_over_l, double *X);

#ifdef __cplusplus
}
#endif
#endif
