// gmi_dense.hip -- exact body-frame preconditioner of the rigid-body solve.
//
// In the body frame the rigid-body system of one swimmer,
//   [ M  -K ] [f]   [r_b]     M: blob-blob mobility (3 nb x 3 nb, SPD)
//   [-K^T 0 ] [U] = [r_c]     K: U = (V, W) -> V + W x ref_j,
// has a fixed matrix: blob positions relative to the body and K depend only
// on the body-frame blob references.  So M^-1 is computed once, kept in HBM
// (3 nb x 3 nb doubles per swimmer: 1.15 GB at 4,002 blobs) with
// G = M^-1 K and S^-1 = (K^T M^-1 K)^-1, and the block inverse
//   w = M^-1 r_b,  U = -S^-1 (r_c + K^T w),  f = w + G U
// serves as the GMRES preconditioner in place of the reference's
// self-mobility diagonal (mobility_solver.cpp:651-727).  GMRES keeps its
// operator and tolerance; the preconditioner only changes how fast it gets
// there.  The reference has no counterpart.
//
// Set-up: M is assembled with the RPY blob-blob forms (cuda_functions.cu:
// 478-517) at the body-frame references, then inverted in place by block
// Gauss-Jordan without pivoting (M is symmetric positive definite): per
// 128-wide block column, the diagonal block is inverted in LDS by one
// workgroup and the rank-128 updates of the whole matrix are rocBLAS DGEMMs.
// Application: one GEMV per swimmer (M^-1 is symmetric, so row i is the
// contiguous column i of the column-major store) between rotations into and
// out of the body frame.
#include "gmi_internal.hpp"
#include "gmi_ctx.hpp"
#include "gmi_lie.hpp"
#include "gmi_sim_state.hpp"


#include <algorithm>
#include <cmath>
#include <vector>

namespace gmi {

constexpr int kSym = 64;  // tile width of the symmetric product
struct SymTile { int sw, I, J, pad; };
struct SymRow { int sw, K, nbl, tile0; };  // block row K of a swimmer with nbl block rows

struct DensePrec {
  std::vector<size_t> moff_h;   // per swimmer: first element of its matrix
  std::vector<int> voff_h;      // per swimmer: first row of its body-frame vector
  int nrows = 0, maxn = 0;      // total rows (sum 3 nb), largest 3 nb
  double *Minv = nullptr;       // per swimmer column-major 3nb x 3nb (set-up only)
  double *Mpk = nullptr;        // lower-triangle tiles of M^-1, kSym x kSym each, contiguous
  double *G = nullptr;          // per swimmer column-major 3nb x 6
  double *Sinv = nullptr;       // [Nsw][36] row-major
  double *t = nullptr, *w = nullptr;  // body-frame blob part of the input, M^-1 t
  double *tc = nullptr;         // body-frame U rows of the input [Nsw][6]
  double *part = nullptr;       // per blob tile: shares of K^T w
  size_t *moff = nullptr;
  int *voff = nullptr;
  // symmetric product over the lower triangle in kSym x kSym tiles
  SymTile *sym_tiles = nullptr;  // (swimmer, block row I >= block column J)
  SymRow *sym_rows = nullptr;    // (swimmer, block row) of the sums
  int nsym_tiles = 0, nsym_rows = 0;
  double *sym_part = nullptr;    // per tile: row partial [kSym] | column partial [kSym]
  std::vector<void *> owned;
};

namespace {

constexpr double kPiD = 3.14159265358979323846264338327950288;
constexpr int kDiag = 128;  // Gauss-Jordan block width

// M at the body-frame references of one swimmer, column-major (element
// (row, col) at row + col * n), rows/cols ordered (component, blob):
// row = c * nb + i.  Thread = blob i, blockIdx.y = blob j.
__global__ __launch_bounds__(64) void k_dense_assemble(const double *__restrict__ bref, int ldb,
                                                       int b0, int nb, double a, double mu,
                                                       double *__restrict__ M) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y;
  if (i >= nb) return;
  const size_t n = size_t(3) * nb;
  const double dx = bref[b0 + i] - bref[b0 + j];
  const double dy = bref[ldb + b0 + i] - bref[ldb + b0 + j];
  const double dz = bref[2 * ldb + b0 + i] - bref[2 * ldb + b0 + j];
  const double r2 = dx * dx + dy * dy + dz * dz;
  const double c8 = 1.0 / (8.0 * kPiD * mu);
  double A, B;
  if (r2 > 4.0 * a * a) {
    const double ri = 1.0 / sqrt(r2), ri2 = ri * ri;
    A = ri * (c8 + c8 * (2.0 * a * a / 3.0) * ri2);
    B = ri2 * ri * (c8 - c8 * 2.0 * a * a * ri2);
  } else {
    const double r = sqrt(r2), c6a = 1.0 / (6.0 * kPiD * mu * a);
    const double b = 3.0 * r / (32.0 * a);
    A = (1.0 - 3.0 * b) * c6a;
    B = r > 0.0 ? b * c6a / r2 : 0.0;
  }
  const double d[3] = {dx, dy, dz};
  for (int c = 0; c < 3; c++)
    for (int e = 0; e < 3; e++)
      M[size_t(c) * nb + i + (size_t(e) * nb + j) * n] = (c == e ? A : 0.0) + B * d[c] * d[e];
}

// Gauss-Jordan inverse (no pivoting: the diagonal blocks of the Schur
// complements of an SPD matrix are SPD) of the kb x kb diagonal block of A
// (column-major, ld n) at (k0, k0) into D (column-major, ld kb).  1024
// threads hold the block in registers: thread t owns row t / 8, columns
// 16 (t % 8) .. + 15 (rows and columns past kb are the identity's).  Each
// pivot step publishes row k and column k in LDS (double-buffered, so one
// barrier per step) and every thread updates its 16 entries.
__global__ __launch_bounds__(1024) void k_dense_diag_inv(const double *__restrict__ A, size_t n,
                                                         int k0, int kb, double *__restrict__ D,
                                                         int *flags) {
  static_assert(kDiag == 128, "thread layout: 128 rows x 8 groups of 16 columns");
  __shared__ double srow[2][kDiag], scol[2][kDiag];
  const int t = threadIdx.x, i = t >> 3, g = t & 7;
  double v[16];
#pragma unroll
  for (int u = 0; u < 16; u++) {
    const int c = 16 * g + u;
    v[u] = (i < kb && c < kb) ? A[size_t(k0 + i) + size_t(k0 + c) * n] : (i == c ? 1.0 : 0.0);
  }
  for (int k = 0; k < kb; k++) {
    const int b = k & 1;
    if (i == k)
#pragma unroll
      for (int u = 0; u < 16; u++) srow[b][16 * g + u] = v[u];
    if (g == (k >> 4))
#pragma unroll
      for (int u = 0; u < 16; u++)
        if (u == (k & 15)) scol[b][i] = v[u];
    __syncthreads();
    const double p = srow[b][k];
    if (!(p > 0.0)) {
      if (t == 0) flags[F_SING] = 1;
      return;
    }
    const double inv = 1.0 / p;
    if (i == k) {
      // row k scaled: L(k, j) = L(k, j) / p, L(k, k) = 1 / p
#pragma unroll
      for (int u = 0; u < 16; u++) {
        const int c = 16 * g + u;
        v[u] = (c == k ? 1.0 : v[u]) * inv;
      }
    } else {
      // other rows: L(i, j) -= L(i, k) L(k, j) / p; column k: L(i, k) = -L(i, k) / p
      const double ci = scol[b][i];
#pragma unroll
      for (int u = 0; u < 16; u++) {
        const int c = 16 * g + u;
        const double rk = (c == k ? 1.0 : srow[b][c]) * inv;
        v[u] = (c == k ? 0.0 : v[u]) - ci * rk;
      }
    }
  }
  if (i < kb)
#pragma unroll
    for (int u = 0; u < 16; u++) {
      const int c = 16 * g + u;
      if (c < kb) D[i + size_t(c) * kb] = v[u];
    }
}

// Gauss-Jordan fix-ups of block column k0: A[K, :] <- T (kb x n, ld kb) and
// A[K, K] <- D.
__global__ void k_dense_fix_rows(double *__restrict__ A, size_t n, int k0, int kb,
                                 const double *__restrict__ T, const double *__restrict__ D) {
  const size_t e = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= size_t(kb) * n) return;
  const int r = int(e % kb);
  const size_t c = e / kb;
  const bool inK = c >= size_t(k0) && c < size_t(k0 + kb);
  A[k0 + r + c * n] = inK ? D[r + (c - k0) * kb] : T[e];
}

// K of one swimmer at its body-frame references, column-major 3nb x 6:
// column e < 3 is e_e on every blob, column 3 + e is e_e x ref_j.
__global__ void k_dense_k(const double *__restrict__ bref, int ldb, int b0, int nb,
                          double *__restrict__ K) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nb) return;
  const size_t n = size_t(3) * nb;
  const double r[3] = {bref[b0 + j], bref[ldb + b0 + j], bref[2 * ldb + b0 + j]};
  for (int e = 0; e < 3; e++)
    for (int c = 0; c < 3; c++) {
      K[size_t(c) * nb + j + size_t(e) * n] = (c == e) ? 1.0 : 0.0;
      // (e_e x r)_c
      const double ex[3] = {e == 0 ? 1.0 : 0.0, e == 1 ? 1.0 : 0.0, e == 2 ? 1.0 : 0.0};
      const double cr[3] = {ex[1] * r[2] - ex[2] * r[1], ex[2] * r[0] - ex[0] * r[2],
                            ex[0] * r[1] - ex[1] * r[0]};
      K[size_t(c) * nb + j + size_t(3 + e) * n] = cr[c];
    }
}

// S^-1 from S (column-major 6x6, symmetric) by Gauss-Jordan with pivoting.
__global__ void k_dense_sinv(const double *__restrict__ S, double *__restrict__ Sinv, int *flags) {
  double A[36];
  int piv[6];
  for (int k = 0; k < 36; k++) A[k] = S[(k % 6) * 6 + k / 6];  // to row-major
  for (int i = 0; i < 6; i++) piv[i] = i;
  for (int k = 0; k < 6; k++) {
    int p = k;
    double best = fabs(A[6 * k + k]);
    for (int i = k + 1; i < 6; i++)
      if (fabs(A[6 * i + k]) > best) { best = fabs(A[6 * i + k]); p = i; }
    if (best == 0.0) { flags[F_SING] = 1; return; }
    if (p != k) {
      for (int j = 0; j < 6; j++) { const double x = A[6 * k + j]; A[6 * k + j] = A[6 * p + j]; A[6 * p + j] = x; }
      const int x = piv[k]; piv[k] = piv[p]; piv[p] = x;
    }
    const double inv = 1.0 / A[6 * k + k];
    A[6 * k + k] = 1.0;
    for (int j = 0; j < 6; j++) A[6 * k + j] *= inv;
    for (int i = 0; i < 6; i++) {
      if (i == k) continue;
      const double g = A[6 * i + k];
      A[6 * i + k] = 0.0;
      for (int j = 0; j < 6; j++) A[6 * i + j] -= g * A[6 * k + j];
    }
  }
  for (int j = 0; j < 6; j++)
    for (int i = 0; i < 6; i++) Sinv[6 * i + piv[j]] = A[6 * i + j];
}

// ---- application ----------------------------------------------------------

// Per blob tile: normalise `in` (norm_slot >= 0) and rotate into the body
// frame: blob rows to t (per swimmer, (component, blob) order), U rows to tc.
__global__ __launch_bounds__(256) void k_dense_in(SimDev a, const BTile *tiles, const int *sw_tile0,
                                                  const int *voff, double *in, const double *scal,
                                                  int norm_slot, double *t, double *tc,
                                                  const int *flags) {
  if (flags && flags[F_DONE]) return;
  const BTile T = tiles[blockIdx.x];
  const int n = T.sw, th = threadIdx.x;
  const m3 Qm = rotm(ga4(a.body + size_t(n) * BREC + BQ));
  const double nrm = norm_slot >= 0 ? scal[norm_slot] : 1.0;
  const int b0 = a.sw_blob0[n], nb = a.sw_nblob[n];
  if (th < T.n) {
    const int m = T.m0 + th;
    d3 v = ld3(in, a.ldb, m);
    if (norm_slot >= 0) {
      v = {v.x / nrm, v.y / nrm, v.z / nrm};
      st3(in, a.ldb, m, v);
    }
    const d3 u = mtv(Qm, v);
    double *tt = t + voff[n] + (m - b0);
    tt[0] = u.x; tt[nb] = u.y; tt[2 * nb] = u.z;
  }
  if (blockIdx.x == sw_tile0[n] && th < 2) {
    double *p = in + 3 * a.ldb + 6 * n + 3 * th;
    d3 v = ga3(p);
    if (norm_slot >= 0) {
      v = {v.x / nrm, v.y / nrm, v.z / nrm};
      sa3(p, v);
    }
    sa3(tc + 6 * n + 3 * th, mtv(Qm, v));
  }
}

// Lower-triangle tile (I, J) of a swimmer's M^-1 copied to contiguous
// storage: element (r, c) of tile b at Mpk[b kSym^2 + c kSym + r], zero
// outside the matrix (one workgroup per tile; set-up only).
__global__ __launch_bounds__(256) void k_dense_pack(const double *__restrict__ Minv,
                                                    const size_t *__restrict__ moff,
                                                    const int *__restrict__ sw_nblob,
                                                    const SymTile *__restrict__ tl,
                                                    double *__restrict__ Mpk) {
  const SymTile T = tl[blockIdx.x];
  const int n = T.sw, rows = 3 * sw_nblob[n];
  const int r0 = T.I * kSym, c0 = T.J * kSym;
  const int nr = min(kSym, rows - r0), nc = min(kSym, rows - c0);
  const double *A = Minv + moff[n];
  double *P = Mpk + size_t(kSym * kSym) * blockIdx.x;
  const int r = threadIdx.x & (kSym - 1), g = threadIdx.x >> 6;
  for (int c = g; c < kSym; c += 4)
    P[c * kSym + r] = (r < nr && c < nc) ? A[size_t(c0 + c) * rows + r0 + r] : 0.0;
}

// w = M^-1 t reading only the lower triangle of each swimmer's symmetric
// inverse, packed tile by tile (k_dense_pack): one workgroup per kSym x kSym
// tile (I >= J) streams its 32 KB, stages the tile in LDS and writes the
// row partial A_IJ t_J and, below the diagonal, the column partial
// A_IJ^T t_I -- every stored element is read from HBM once for both uses.
// k_dense_symv_sum adds the partials of a block row in a fixed order.
__global__ __launch_bounds__(256) void k_dense_symv_tiles(const double *__restrict__ Mpk,
                                                          const int *__restrict__ voff,
                                                          const int *__restrict__ sw_nblob,
                                                          const SymTile *__restrict__ tl,
                                                          const double *__restrict__ t,
                                                          double *__restrict__ part,
                                                          const int *flags) {
  if (flags && flags[F_DONE]) return;
  constexpr int LD = kSym + 1;  // padded column stride: conflict-free row and column reads
  __shared__ double L[kSym * LD];
  __shared__ double xi[kSym], xj[kSym];
  __shared__ double red[4][kSym];
  const SymTile T = tl[blockIdx.x];
  const int n = T.sw, rows = 3 * sw_nblob[n];
  const int r0 = T.I * kSym, c0 = T.J * kSym;
  const int nr = min(kSym, rows - r0), nc = min(kSym, rows - c0);
  const double *A = Mpk + size_t(kSym * kSym) * blockIdx.x;
  const double *tv = t + voff[n];
  const int th = threadIdx.x, r = th & (kSym - 1), g = th >> 6;
  // load k of the workgroup covers elements [256 k, 256 k + 256) of the
  // tile; all loads of a thread are issued before its LDS stores, so each
  // lane keeps kSym / 4 loads in flight
  constexpr int NL = kSym / 4;
  double v[NL];
#pragma unroll
  for (int k = 0; k < NL; k++) v[k] = A[(g + 4 * k) * kSym + r];
  const int xk = th & (kSym - 1);
  const bool xin = th < kSym ? xk < nr : xk < nc;
  const double xv = tv[xin ? (th < kSym ? r0 : c0) + xk : r0];
#pragma unroll
  for (int k = 0; k < NL; k++) L[(g + 4 * k) * LD + r] = v[k];
  if (th < kSym) xi[xk] = xin ? xv : 0.0;
  else if (th < 2 * kSym) xj[xk] = xin ? xv : 0.0;
  __syncthreads();
  double s = 0.0;
  if (g < 2) {                 // row partial, columns g*32 .. g*32+31
    for (int c = g * 32; c < g * 32 + 32; c++) s += L[c * LD + r] * xj[c];
  } else if (T.I != T.J) {     // column partial, rows (g-2)*32 .. +31
    const int h = (g - 2) * 32;
    for (int q = h; q < h + 32; q++) s += L[r * LD + q] * xi[q];
  }
  red[g][r] = s;
  __syncthreads();
  double *P = part + size_t(2 * kSym) * blockIdx.x;
  if (th < kSym) P[th] = red[0][th] + red[1][th];
  else if (th < 2 * kSym) P[th] = red[2][th - kSym] + red[3][th - kSym];
}

// w rows of block row K: sum over J <= K of the row partials of tiles (K, J)
// and over I > K of the column partials of tiles (I, K); four lane groups
// take every fourth term, combined in group order.
__global__ __launch_bounds__(256) void k_dense_symv_sum(const SymRow *__restrict__ rl,
                                                        const int *__restrict__ voff,
                                                        const int *__restrict__ sw_nblob,
                                                        const double *__restrict__ part,
                                                        double *__restrict__ w, const int *flags) {
  if (flags && flags[F_DONE]) return;
  __shared__ double red[4][kSym];
  const SymRow R = rl[blockIdx.x];
  const int th = threadIdx.x, k = th & (kSym - 1), g = th >> 6;
  // term q: tile (K, q) row partial for q <= K, tile (q, K) column partial
  // above; four independent running sums per lane keep four loads in flight
  auto term = [&](int q) -> double {
    const size_t tile = q <= R.K ? size_t(R.tile0) + size_t(R.K) * (R.K + 1) / 2 + q
                                 : size_t(R.tile0) + size_t(q) * (q + 1) / 2 + R.K;
    return part[size_t(2 * kSym) * tile + (q <= R.K ? 0 : kSym) + k];
  };
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int q = g;
  for (; q + 12 < R.nbl; q += 16) {
    s0 += term(q);
    s1 += term(q + 4);
    s2 += term(q + 8);
    s3 += term(q + 12);
  }
  for (; q < R.nbl; q += 4) s0 += term(q);
  red[g][k] = (s0 + s1) + (s2 + s3);
  __syncthreads();
  const int row = R.K * kSym + th;
  if (th < kSym && row < 3 * sw_nblob[R.sw])
    w[voff[R.sw] + row] = ((red[0][th] + red[1][th]) + red[2][th]) + red[3][th];
}

// Per blob tile: shares of K^T w = (sum w_j, sum ref_j x w_j).
__global__ __launch_bounds__(256) void k_dense_kt(SimDev a, const BTile *tiles, const int *voff,
                                                  const double *w, double *part,
                                                  const int *flags) {
  if (flags && flags[F_DONE]) return;
  __shared__ double sh[4 * 6];
  const BTile T = tiles[blockIdx.x];
  const int n = T.sw, th = threadIdx.x;
  const int b0 = a.sw_blob0[n], nb = a.sw_nblob[n];
  double c[6] = {0, 0, 0, 0, 0, 0};
  if (th < T.n) {
    const int m = T.m0 + th;
    const double *wv = w + voff[n] + (m - b0);
    const d3 v = {wv[0], wv[nb], wv[2 * nb]};
    const d3 cr = cross3(ld3(a.blob_ref, a.ldb, m), v);
    c[0] = v.x; c[1] = v.y; c[2] = v.z; c[3] = cr.x; c[4] = cr.y; c[5] = cr.z;
  }
  const int lane = th & 63, wv = th >> 6;
  for (int k = 0; k < 6; k++) {
    const double r = wave_sum(c[k]);
    if (lane == 0) sh[4 * k + wv] = r;
  }
  __syncthreads();
  if (th < 6) part[6 * blockIdx.x + th] = ((sh[4 * th] + sh[4 * th + 1]) + sh[4 * th + 2]) + sh[4 * th + 3];
}

// Per blob tile: U = -S^-1 (r_c + K^T w) from the swimmer's shares (tile
// order), f = w + G U, and out = R f (blob rows), R U (U rows, first tile).
__global__ __launch_bounds__(256) void k_dense_out(SimDev a, const BTile *tiles, const int *sw_tile0,
                                                   const int *sw_ntile, const int *voff,
                                                   const double *w, const double *G,
                                                   const double *Sinv, const double *tc,
                                                   const double *part, double *out,
                                                   const int *flags) {
  if (flags && flags[F_DONE]) return;
  __shared__ double sU[6];
  const BTile T = tiles[blockIdx.x];
  const int n = T.sw, th = threadIdx.x;
  if (th < 6) {
    double s = 0.0;
    for (int q = sw_tile0[n]; q < sw_tile0[n] + sw_ntile[n]; q++) s += part[6 * q + th];
    sU[th] = tc[6 * n + th] + s;
  }
  __syncthreads();
  double U[6];
  for (int i = 0; i < 6; i++) {
    double s = 0.0;
    for (int k = 0; k < 6; k++) s += Sinv[36 * n + 6 * i + k] * sU[k];
    U[i] = -s;
  }
  const m3 Qm = rotm(ga4(a.body + size_t(n) * BREC + BQ));
  const int b0 = a.sw_blob0[n], nb = a.sw_nblob[n];
  const size_t rows = size_t(3) * nb;
  if (th < T.n) {
    const int m = T.m0 + th;
    const int jl = m - b0;
    const double *g = G + size_t(6) * voff[n];
    double f[3];
    for (int c = 0; c < 3; c++) {
      const size_t row = size_t(c) * nb + jl;
      double s = w[voff[n] + row];
      for (int e = 0; e < 6; e++) s += g[row + e * rows] * U[e];
      f[c] = s;
    }
    st3(out, a.ldb, m, mv(Qm, d3{f[0], f[1], f[2]}));
  }
  if (blockIdx.x == sw_tile0[n] && th < 2) {
    const d3 u = {U[3 * th], U[3 * th + 1], U[3 * th + 2]};
    sa3(out + 3 * a.ldb + 6 * n + 3 * th, mv(Qm, u));
  }
}

inline int nb_of(int n, int b) { return (n + b - 1) / b; }

}  // namespace

void dense_release(SimHost *s) {
  DensePrec *D = s->dense;
  if (!D) return;
  for (void *p : D->owned) (void)hipFree(p);
  delete D;
  s->dense = nullptr;
  s->Zk = nullptr;  // owned by D
}

static int dmalloc(gmi_ctx *c, DensePrec *D, void **p, size_t bytes) {
  int rc = hip_check(c, hipMalloc(p, std::max<size_t>(bytes, 8)), "hipMalloc dense preconditioner");
  if (!rc) D->owned.push_back(*p);
  return rc;
}


int dense_setup(gmi_ctx *c, SimHost *s) {
  const SimDev &d = s->d;
  DensePrec *D = new DensePrec();
  s->dense = D;
  int rc;
  auto bail = [&](int code) {
    dense_release(s);
    return code;
  };
  D->moff_h.resize(d.Nsw);
  D->voff_h.resize(d.Nsw);
  size_t mtot = 0;
  int vtot = 0, maxn = 0;
  for (int n = 0; n < d.Nsw; n++) {
    const int rows = 3 * c->nblob[n];
    D->moff_h[n] = mtot;
    D->voff_h[n] = vtot;
    mtot += size_t(rows) * rows;
    vtot += rows;
    maxn = std::max(maxn, rows);
  }
  D->nrows = vtot;
  D->maxn = maxn;
  if ((rc = dmalloc(c, D, (void **)&D->Minv, sizeof(double) * mtot)) ||
      (rc = dmalloc(c, D, (void **)&D->G, sizeof(double) * 6 * size_t(vtot))) ||
      (rc = dmalloc(c, D, (void **)&D->Sinv, sizeof(double) * 36 * d.Nsw)) ||
      (rc = dmalloc(c, D, (void **)&D->t, sizeof(double) * vtot)) ||
      (rc = dmalloc(c, D, (void **)&D->w, sizeof(double) * vtot)) ||
      (rc = dmalloc(c, D, (void **)&D->tc, sizeof(double) * 6 * d.Nsw)) ||
      (rc = dmalloc(c, D, (void **)&D->part, sizeof(double) * 6 * std::max(1, s->nbtiles))) ||
      (rc = dmalloc(c, D, (void **)&D->moff, sizeof(size_t) * d.Nsw)) ||
      (rc = dmalloc(c, D, (void **)&D->voff, sizeof(int) * d.Nsw)) ||
      // the preconditioned Arnoldi directions: each solve applies M^-1 once
      // per step and not again for the update
      (rc = dmalloc(c, D, (void **)&s->Zk, sizeof(double) * size_t(d.npad) * (d.maxit + 1))) ||
      (rc = hip_check(c, hipMemsetAsync(s->Zk, 0, sizeof(double) * size_t(d.npad) * (d.maxit + 1),
                                        c->stream), "zero directions")))
    return bail(rc);
  // Gauss-Jordan workspaces of the largest swimmer
  double *P = nullptr, *T = nullptr, *Dd = nullptr, *K = nullptr, *S = nullptr;
  if ((rc = dmalloc(c, D, (void **)&P, sizeof(double) * size_t(maxn) * kDiag)) ||
      (rc = dmalloc(c, D, (void **)&T, sizeof(double) * size_t(maxn) * kDiag)) ||
      (rc = dmalloc(c, D, (void **)&Dd, sizeof(double) * kDiag * kDiag)) ||
      (rc = dmalloc(c, D, (void **)&K, sizeof(double) * size_t(maxn) * 6)) ||
      (rc = dmalloc(c, D, (void **)&S, sizeof(double) * 36)))
    return bail(rc);
  // tiles of the symmetric product, per swimmer in (I, J <= I) order
  std::vector<SymTile> tl;
  std::vector<SymRow> rl;
  for (int n = 0; n < d.Nsw; n++) {
    const int nbl = nb_of(3 * c->nblob[n], kSym), t0 = int(tl.size());
    for (int I = 0; I < nbl; I++) {
      rl.push_back({n, I, nbl, t0});
      for (int J = 0; J <= I; J++) tl.push_back({n, I, J, 0});
    }
  }
  D->nsym_tiles = int(tl.size());
  D->nsym_rows = int(rl.size());
  if ((rc = dmalloc(c, D, (void **)&D->sym_tiles, sizeof(SymTile) * tl.size())) ||
      (rc = dmalloc(c, D, (void **)&D->sym_rows, sizeof(SymRow) * rl.size())) ||
      (rc = dmalloc(c, D, (void **)&D->sym_part, sizeof(double) * 2 * kSym * tl.size())) ||
      (rc = hip_check(c, hipMemcpyAsync(D->sym_tiles, tl.data(), sizeof(SymTile) * tl.size(),
                                        hipMemcpyHostToDevice, c->stream), "upload")) ||
      (rc = hip_check(c, hipMemcpyAsync(D->sym_rows, rl.data(), sizeof(SymRow) * rl.size(),
                                        hipMemcpyHostToDevice, c->stream), "upload")) ||
      (rc = hip_check(c, hipStreamSynchronize(c->stream), "sync")))  // tl, rl are locals
    return bail(rc);
  if ((rc = hip_check(c, hipMemcpyAsync(D->moff, D->moff_h.data(), sizeof(size_t) * d.Nsw,
                                        hipMemcpyHostToDevice, c->stream), "upload")) ||
      (rc = hip_check(c, hipMemcpyAsync(D->voff, D->voff_h.data(), sizeof(int) * d.Nsw,
                                        hipMemcpyHostToDevice, c->stream), "upload")))
    return bail(rc);
  hipLaunchKernelGGL(k_set_ints, dim3(1), dim3(64), 0, c->stream, d.flags + F_SING, 1, 0,
                     nullptr);
  for (int n = 0; n < d.Nsw; n++) {
    const int nb = c->nblob[n], rows = 3 * nb;
    const size_t ld = size_t(rows);
    double *A = D->Minv + D->moff_h[n];
    hipLaunchKernelGGL(k_dense_assemble, dim3(nb_of(nb, 64), nb), dim3(64), 0, c->stream,
                       d.blob_ref, d.ldb, c->blob_off[n], nb, d.ab, d.mu, A);
    if ((rc = launch_ok(c, "assemble"))) return bail(rc);
    for (int k0 = 0; k0 < rows; k0 += kDiag) {
      const int kb = std::min(kDiag, rows - k0);
      hipLaunchKernelGGL(k_dense_diag_inv, dim3(1), dim3(1024), 0, c->stream, A, ld, k0, kb, Dd,
                         d.flags);
      // T = D A[K, :]  (kb x rows)
      if ((rc = hip_check(c, launch_dgemm(false, kb, rows, kb, 1.0, Dd, kb, A + k0, ld, 0.0, T,
                                          kb, c->stream), "dgemm T")))
        return bail(rc);
      // P = A[:, K]
      if ((rc = hip_check(c, hipMemcpyAsync(P, A + size_t(k0) * ld, sizeof(double) * ld * kb,
                                            hipMemcpyDeviceToDevice, c->stream), "copy P")))
        return bail(rc);
      // A -= P T
      if ((rc = hip_check(c, launch_dgemm(false, rows, rows, kb, -1.0, P, ld, T, kb, 1.0, A, ld,
                                          c->stream), "dgemm update")))
        return bail(rc);
      // A[:, K] = -P D, then A[K, :] = T and A[K, K] = D
      if ((rc = hip_check(c, launch_dgemm(false, rows, kb, kb, -1.0, P, ld, Dd, kb, 0.0,
                                          A + size_t(k0) * ld, ld, c->stream), "dgemm column")))
        return bail(rc);
      hipLaunchKernelGGL(k_dense_fix_rows, dim3(nb_of(kb * rows, 256)), dim3(256), 0, c->stream,
                         A, ld, k0, kb, T, Dd);
      if ((rc = launch_ok(c, "gauss-jordan"))) return bail(rc);
    }
    // G = M^-1 K, S = K^T G
    hipLaunchKernelGGL(k_dense_k, dim3(nb_of(nb, 256)), dim3(256), 0, c->stream, d.blob_ref,
                       d.ldb, c->blob_off[n], nb, K);
    double *Gn = D->G + size_t(6) * D->voff_h[n];
    if ((rc = hip_check(c, launch_dgemm(false, rows, 6, rows, 1.0, A, ld, K, ld, 0.0, Gn, ld,
                                        c->stream), "dgemm G")) ||
        (rc = hip_check(c, launch_dgemm(true, 6, 6, rows, 1.0, K, ld, Gn, ld, 0.0, S, 6,
                                        c->stream), "dgemm S")))
      return bail(rc);
    hipLaunchKernelGGL(k_dense_sinv, dim3(1), dim3(1), 0, c->stream, S, D->Sinv + 36 * n, d.flags);
    if ((rc = launch_ok(c, "schur"))) return bail(rc);
  }
  int sing = 0;
  if ((rc = hip_check(c, hipMemcpyAsync(&sing, d.flags + F_SING, sizeof(int), hipMemcpyDeviceToHost,
                                        c->stream), "flags")) ||
      (rc = hip_check(c, hipStreamSynchronize(c->stream), "sync")))
    return bail(rc);
  if (sing) return bail(fail(c, GMI_ERR_SINGULAR, "blob mobility not positive definite"));
  // the product reads the lower triangle packed tile by tile; the full
  // matrices are not needed after this (half the HBM stays resident)
  if ((rc = dmalloc(c, D, (void **)&D->Mpk,
                    sizeof(double) * size_t(kSym * kSym) * size_t(D->nsym_tiles))))
    return bail(rc);
  hipLaunchKernelGGL(k_dense_pack, dim3(D->nsym_tiles), dim3(256), 0, c->stream, D->Minv, D->moff,
                     d.sw_nblob, D->sym_tiles, D->Mpk);
  if ((rc = launch_ok(c, "pack")) ||
      (rc = hip_check(c, hipStreamSynchronize(c->stream), "sync")))
    return bail(rc);
  D->owned.erase(std::find(D->owned.begin(), D->owned.end(), static_cast<void *>(D->Minv)));
  if ((rc = hip_check(c, hipFree(D->Minv), "hipFree full inverse"))) return bail(rc);
  D->Minv = nullptr;
  return GMI_OK;
}

int dense_apply(gmi_ctx *c, SimHost *s, double *in, double *out, const int *skip, int norm_slot) {
  const SimDev &d = s->d;
  DensePrec *D = s->dense;
  hipLaunchKernelGGL(k_dense_in, dim3(s->nbtiles), dim3(256), 0, c->stream, d, s->btiles,
                     s->sw_tile0, D->voff, in, s->scal, norm_slot, D->t, D->tc, skip);
  hipLaunchKernelGGL(k_dense_symv_tiles, dim3(D->nsym_tiles), dim3(256), 0, c->stream, D->Mpk,
                     D->voff, d.sw_nblob, D->sym_tiles, D->t, D->sym_part, skip);
  hipLaunchKernelGGL(k_dense_symv_sum, dim3(D->nsym_rows), dim3(256), 0, c->stream, D->sym_rows,
                     D->voff, d.sw_nblob, D->sym_part, D->w, skip);
  hipLaunchKernelGGL(k_dense_kt, dim3(s->nbtiles), dim3(256), 0, c->stream, d, s->btiles, D->voff,
                     D->w, D->part, skip);
  hipLaunchKernelGGL(k_dense_out, dim3(s->nbtiles), dim3(256), 0, c->stream, d, s->btiles,
                     s->sw_tile0, s->sw_ntile, D->voff, D->w, D->G, D->Sinv, D->tc, D->part, out,
                     skip);
  return hip_check(c, hipGetLastError(), "dense preconditioner");
}

}  // namespace gmi
