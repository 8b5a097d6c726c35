// gmi_sim_state.hpp -- device state of the implicit time step (not public).
#pragma once
#include "gmi_internal.hpp"
#include <map>
#include <utility>
#include "../../include/gmi_sim.h"
#include "gmi_krylov.hpp"

namespace gmi {

// body record (per swimmer), doubles
enum {
  BX = 0, BXM1 = 3, BXM2 = 6, BU = 9, BUM1 = 12, BQ = 15, BQM1 = 19, BQI = 23,
  BF = 32, BN = 41, BK = 77, BS = 113, BREC = 152
};

struct SimDev {
  // sizes
  int S, Nf, Nsw, Ns, Nb, lds, ldb, nbroy, npad, maxit;
  double DL, mu, as, ab, kb, kt, dt, tol, gamma, gtol;
  int neuler;
  double camp, cper;
  double efm;          // END_FORCE_MAGNITUDE (0: no end force)
  const double *pert;  // [Nf][6] perturbations at nt = 0, or null
  // index maps
  const int *sw_fil0, *sw_nfil, *sw_blob0, *sw_nblob, *fil_sw, *blob_sw;
  const double *sw_R;  // body radius per swimmer (Schur RPY correction)
  // segment SoA
  double *X, *Xm1, *Xm2, *Q, *Qm1, *U, *Um1, *Fint, *Fs, *Vs;
  // filament
  double *fil_ref, *phase, *lam, *lamm1, *lamm2;  // lam: [Nf][3S]
  double *teth, *clamp;                           // [Nf][9]: value, m1, m2
  double *fil_bf;                                 // [Nf][6]
  // body
  double *body;                                   // [Nsw][BREC]
  double *blob_ref;                               // SoA [3][ldb]
  double *Xb, *Vb, *Frep;
  // Jacobian
  double *Jw, *Jinv, *ecb, *Bblk, *fil_BA, *fil_b2;
  int *flags;
  int *gj_redo;  // [Nf] 1: the pivot-free inverse failed its check (pivoted rerun)
  // the step being solved, set once per step on the device so that captured
  // Broyden iterations serve every step: step_t[0] = (nt + 1) dt, step_nt[0] = nt
  double *step_t;
  int *step_nt;
};


struct Learn;      // learned warm start (gmi_learn.hip)
struct DensePrec;  // exact body-frame preconditioner (gmi_dense.hip)

// up to 256 consecutive blobs of one swimmer
struct BTile {
  int sw, m0, n;
};

// GMRES state (x, rhs, basis, flags, graphs) comes from KrylovState.
struct SimHost : KrylovState {
  SimDev d;
  gmi_sim_params p;          // pointer members re-pointed at the copies below
  std::vector<double> blob_refs_h, fil_refs_h;
  std::vector<int> nfil;
  // device buffers owned
  std::vector<void *> owned;
  double *rot = nullptr;  // rotated N_ref and (K^T K)^-1 per swimmer [Nsw][72]
  double *err = nullptr, *nerr = nullptr, *upd = nullptr, *C = nullptr, *D = nullptr,
         *bd = nullptr;
  double *err_first = nullptr;  // the buffer err pointed at first (err and nerr swap)
  int iter = 0;
  int captured_iters = 0;  // Broyden iterations run from captured graphs
  int precision = GMI_FP64;
  int last_gmres = 0;
  double iters_avg = 0.0;  // moving average of Broyden iterations per step
  // history-fitted GMRES initial guess (0: the previous solution); the
  // history's (count, head) live on the device so that captured iterations
  // can push into it
  int hist_m = 0;
  int *hist_st = nullptr;
  double *hist_b = nullptr, *hist_x = nullptr, *hist_bnew = nullptr, *hist_G = nullptr,
         *hist_Gw = nullptr, *hist_part = nullptr, *hist_coef = nullptr, *hist_gsave = nullptr;
  bool gj_lds = false;     // Gauss-Jordan inverse with the matrix in LDS
  bool gj_pivot_always = false;  // skip the checked pivot-free inverse
  Learn *learn = nullptr;
  DensePrec *dense = nullptr;  // null: the reference's self-mobility preconditioner
  // blob tiles for the preconditioner and K products
  BTile *btiles = nullptr;
  int nbtiles = 0;
  int *sw_tile0 = nullptr, *sw_ntile = nullptr;
  double *tpart = nullptr, *zpart = nullptr;  // tile shares of -K^T v (preconditioner, K products)
  // chunks of up to 64 blobs of one swimmer (fused Arnoldi steps), the blob
  // tile each lies in, and their shares of -K^T v
  BTile *chunks = nullptr;
  int nchunks = 0;
  int *ch_tile = nullptr, *sw_chunk0 = nullptr, *sw_nchunk = nullptr;
  double *ch_kpart = nullptr, *ch_zpart = nullptr;
};


// learned warm start hooks (gmi_learn.hip)
void learn_release(SimHost *s);
int learn_after_initial_guess(gmi_ctx *c, SimHost *s, int nt, bool *applied);
int learn_after_convergence(gmi_ctx *c, SimHost *s, int nt);
double *learn_correction(SimHost *s);
// exact body-frame preconditioner (gmi_dense.hip)
int dense_setup(gmi_ctx *c, SimHost *s);
int dense_apply(gmi_ctx *c, SimHost *s, double *in, double *out, const int *skip, int norm_slot);
void dense_release(SimHost *s);
int learn_back_to_guess(gmi_ctx *c, SimHost *s, double **delta);

}  // namespace gmi
