// gmi_krylov.hpp -- device GMRES shared by the rigid-body solve
// (gmi_sim.hip) and the prescribed-cilia saddle-point solve (gmi_cilia.hip).
//
// Right-preconditioned GMRES (mobility_solver.cpp:970-1354) with the Krylov
// basis, Hessenberg matrix and Givens rotations resident in HBM.  Arnoldi
// uses classical Gram-Schmidt, one pass per step by default (as the
// reference's one modified Gram-Schmidt pass) or two (reorthogonalised):
// a pass is one batched dot-product sweep over all previous columns instead
// of k dependent ones.  The host only reads the convergence flag once
// per chunk of Arnoldi steps; a chunk can be captured once as a HIP graph
// and replayed by later solves.
//
// An operator type Op supplies
//   int mult(double *z, double *out, int mode, const int *done, bool zpart_ready)
//       out = A z; mode 0: FP64 pairs, 1: FP32 pairs, 2: both launched, each
//       gated by its device flag (F_SKIP64 / F_SKIP32); zpart_ready: z came
//       straight from precond() (which may leave per-tile shares behind);
//   int precond(double *in, double *out, const int *skip, int norm_slot)
//       out = P^-1 in; norm_slot >= 0: `in` is first divided in place by
//       scal[norm_slot] (the previous Arnoldi vector's norm).
//   bool fused(), int fused_parts(),
//   int fused_step(KrylovState *s, int it, bool chunk_start, bool mixed,
//                  const GivensArgs &g)
//       an operator may run a whole Arnoldi step itself (preconditioner,
//       operator, both Gram-Schmidt passes) on a row partition of its own,
//       folding the Givens step of iteration it-1 into its first kernel
//       unless the step starts a chunk (then a k_givens launch ran it); it
//       leaves the new vector unnormalised, fused_parts() norm shares in
//       s->part, and h1, h2 in s->h1, s->h2.
// Every kernel either operator launches must return at once when `done`
// (the F_DONE flag) is set, so that a chunk needs no host decisions.
#pragma once
#include "gmi_ctx.hpp"
#include "gmi_wave.hpp"
#include "../../include/gmi.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <utility>

namespace gmi {

// device flags
// F_HOLD: set while a captured Broyden iteration's GMRES is unfinished; the
// kernels after the solve in that graph are no-ops then (the host finishes
// the iteration with eager launches).
enum { F_DONE = 0, F_ITS = 1, F_BIG = 2, F_NAN = 3, F_SING = 4, F_SKIP64 = 5, F_SKIP32 = 6,
       F_HOLD = 7, F_N = 8 };

// A captured stream segment: the graph is kept as long as its executable
// (both are destroyed together).
struct CapturedGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  void destroy() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
};

// Captures what `enqueue()` launches on the context stream into *out.  On
// any failure the capture is ended, everything it created is destroyed and
// *out is left empty.
template <class F>
int capture_graph(gmi_ctx *c, F &&enqueue, CapturedGraph *out) {
  int rc;
  if ((rc = hip_check(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal),
                      "begin capture")))
    return rc;
  const int erc = enqueue();
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(c->stream, &graph);
  if (erc || ec != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipGetLastError();
    return erc ? erc : hip_check(c, ec, "end capture");
  }
  hipGraphExec_t exec = nullptr;
  const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (ei != hipSuccess) {
    (void)hipGraphDestroy(graph);
    return hip_check(c, ei, "instantiate");
  }
  out->graph = graph;
  out->exec = exec;
  return GMI_OK;
}

// Device buffers and host state of one GMRES solver.
struct KrylovState {
  double *x = nullptr, *rhs = nullptr, *Qk = nullptr, *H = nullptr, *cs = nullptr,
         *sn = nullptr, *beta = nullptr, *h1 = nullptr, *h2 = nullptr, *y = nullptr,
         *z = nullptr, *tmp = nullptr, *scal = nullptr, *part = nullptr, *pp = nullptr;
  size_t pp_half = 0;     // column dot partials: pp[0, pp_half) and pp[pp_half, 2 pp_half)
  // Optional: the preconditioned directions z_k = P^-1 q_k, one per Arnoldi
  // step (same layout as Qk).  When set, the solution update is x += Z y and
  // the preconditioner is not applied again after the last step (the update
  // of flexible GMRES; identical in exact arithmetic for a fixed P).
  double *Zk = nullptr;
  // Optional: the right-hand side of the least-squares problem before each
  // Givens rotation touches it (bpre[c] = beta[c] as it was when column c's
  // rotation was computed).  With it the rotation of a column can be
  // computed redundantly by every workgroup of a kernel while one of them
  // writes beta (the fused Arnoldi step).
  double *bpre = nullptr;
  int *flags = nullptr;   // device flags (F_*)
  int *h_flags = nullptr; // pinned copies
  double *h_scal = nullptr;
  int cols = 0;           // Arnoldi steps the basis holds (H is (cols+1) x (cols+1))
  int chunk = 6;          // Arnoldi steps per host check
  int gs_passes = 1;      // Gram-Schmidt passes per Arnoldi step (1, or 2: reorthogonalised)
  bool use_graphs = true;
  int graph_gen = -1;
  // captured Arnoldi chunks, keyed by (mixed, first iteration)
  std::map<std::pair<int, int>, CapturedGraph> graphs;
  // captured outer iterations of a solver built on this one (their key is
  // the owner's)
  std::map<long long, CapturedGraph> outer_graphs;
  void drop_graphs() {
    for (auto &kv : graphs) kv.second.destroy();
    graphs.clear();
    for (auto &kv : outer_graphs) kv.second.destroy();
    outer_graphs.clear();
  }
};

namespace {


// Sum over the NT threads of a workgroup, in every thread: wave sums, then
// a fixed pairwise tree over the NT / 64 wave sums (two barriers; none for
// a single wave).  sh: NT / 64 doubles.
template <int NT>
__device__ double block_sum(double v, double *sh) {
  static_assert(NT % 64 == 0 && NT <= 1024, "whole waves");
  constexpr int NW = NT / 64;
  const double w = wave_sum(v);
  if (NW == 1) return w;
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = w;
  __syncthreads();
  double p[NW];
#pragma unroll
  for (int i = 0; i < NW; i++) p[i] = sh[i];
#pragma unroll
  for (int h = 1; h < NW; h *= 2)
#pragma unroll
    for (int i = 0; i + h < NW; i += 2 * h) p[i] += p[i + h];
  __syncthreads();
  return p[0];
}

// Dot products with many columns, chunked over the rows: workgroup b takes
// R consecutive rows (R a multiple of 256, at most 128 workgroups until R
// reaches kMaxRows) and writes one partial per column, part[j * nb + b];
// each consumer sums a column's partials in workgroup order.  No
// cross-workgroup synchronisation inside a kernel: the kernel boundary makes
// the partials visible.
constexpr int kMaxDotBlocks = 128;
constexpr int kMaxRows = 6144;     // rows per workgroup (LDS: R doubles, + kMaxCols)
constexpr int kMaxCols = 1024;     // Krylov / Broyden columns held in LDS

// partials of rows [e0, e0+R) (values in sw) with columns j < k; wave w
// takes the columns j = w (mod 4)
// Columns go in batches of kDotBatch per wave so that each lane keeps that
// many independent loads in flight (a column's sum order is unchanged: lane
// partials over its rows, then the wave sum).
constexpr int kDotBatch = 8;
__device__ void col_dots(const double *Qb, size_t ldq, int k, const double *sw, int e0, int n,
                         int R, double *part) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int rmax = min(R, n - e0);
  for (int j0 = wave; j0 < k; j0 += 4 * kDotBatch) {
    double acc[kDotBatch];
    const double *q[kDotBatch];
#pragma unroll
    for (int c = 0; c < kDotBatch; c++) {
      acc[c] = 0.0;
      // past the last column a batch re-reads column k-1 and stores nothing
      q[c] = Qb + ldq * min(j0 + 4 * c, k - 1) + e0;
    }
    for (int r = lane; r < rmax; r += 64) {
      const double s = sw[r];
#pragma unroll
      for (int c = 0; c < kDotBatch; c++) acc[c] += q[c][r] * s;
    }
#pragma unroll
    for (int c = 0; c < kDotBatch; c++) {
      const double a = wave_sum(acc[c]);
      const int j = j0 + 4 * c;
      if (lane == 0 && j < k) part[size_t(j) * gridDim.x + blockIdx.x] = a;
    }
  }
}

// h[j] = sum over workgroups of part[j][b] (workgroup order), j < k, into
// LDS (and, from workgroup 0, into global hout)
__device__ void col_sums(const double *part, int k, int nb, double *hs, double *hout) {
  for (int j = threadIdx.x; j < k; j += blockDim.x) {
    const double *p = part + size_t(j) * nb;
    double s = 0.0;
    // all the loads of a column in flight at once; the sum keeps its order
#pragma unroll 16
    for (int b = 0; b < nb; b++) s += p[b];
    hs[j] = s;
    if (hout && blockIdx.x == 0) hout[j] = s;
  }
}

// partials of h[j] = <Q_j, w>, j < k
__global__ __launch_bounds__(256) void k_pdots(const double *Qb, size_t ldq, int k,
                                               const double *w, int n, int R, double *part,
                                               const int *flags) {
  if (flags && flags[F_DONE]) return;
  extern __shared__ double sw[];
  const int e0 = blockIdx.x * R;
  for (int r = threadIdx.x; r < R; r += 256) sw[r] = (e0 + r < n) ? w[e0 + r] : 0.0;
  __syncthreads();
  col_dots(Qb, ldq, k, sw, e0, n, R, part);
}

// h[j] = sum of the partials, one workgroup per column (large n: many
// partials per column)
__global__ __launch_bounds__(256) void k_colsum_wide(const double *part, int nb, double *h,
                                                     const int *flags) {
  if (flags && flags[F_DONE]) return;
  __shared__ double sh[256];
  const double *p = part + size_t(blockIdx.x) * nb;
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) s += p[b];
  s = block_sum<256>(s, sh);
  if (threadIdx.x == 0) h[blockIdx.x] = s;
}

// h[j] = sum of the partials (one workgroup per 256 columns)
__global__ __launch_bounds__(256) void k_colsum(const double *part, int k, int nb, double *h,
                                                const int *skip) {
  if (skip && *skip) return;
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= k) return;
  const double *p = part + size_t(j) * nb;
  double s = 0.0;
#pragma unroll 4
  for (int b = 0; b < nb; b++) s += p[b];
  h[j] = s;
}

// Gram-Schmidt pass 1: h1 = sum of the pass-1 partials; w -= sum_j h1[j] Q_j;
// then the partials of the pass-2 dots <Q_j, w>
// (h_ready: h1 was summed by k_colsum_wide already)
__global__ __launch_bounds__(256) void k_psub_dots(const double *Qb, size_t ldq, int k,
                                                   const double *part_in, double *h1, double *w,
                                                   int n, int R, double *part_out,
                                                   const int *flags, int h_ready) {
  if (flags[F_DONE]) return;
  __shared__ double hs[kMaxCols];
  extern __shared__ double sw[];
  if (h_ready)
    for (int j = threadIdx.x; j < k; j += 256) hs[j] = h1[j];
  else
    col_sums(part_in, k, gridDim.x, hs, h1);
  __syncthreads();
  const int e0 = blockIdx.x * R;
  for (int r = threadIdx.x; r < R; r += 256) {
    const int e = e0 + r;
    double v = 0.0;
    if (e < n) {
      v = w[e];
#pragma unroll 8
      for (int j = 0; j < k; j++) v -= hs[j] * Qb[ldq * j + e];
      w[e] = v;
    }
    sw[r] = v;
  }
  __syncthreads();
  col_dots(Qb, ldq, k, sw, e0, n, R, part_out);
}

// Gram-Schmidt pass 2: h2 = sum of the pass-2 partials; w -= sum_j h2[j] Q_j;
// the workgroup's share of |w|^2.  With one pass per step it follows the
// first dots directly (h2 is then the only pass's h) and workgroup 0 zeroes
// hzero, the unused pass's coefficients that the Givens step adds.
__global__ __launch_bounds__(256) void k_psub_norm(const double *Qb, size_t ldq, int k,
                                                   const double *part_in, double *h2, double *w,
                                                   int n, int R, double *norm_part,
                                                   const int *flags, int h_ready,
                                                   double *hzero) {
  if (flags[F_DONE]) return;
  if (hzero && blockIdx.x == 0)
    for (int j = threadIdx.x; j < k; j += 256) hzero[j] = 0.0;
  __shared__ double hs[kMaxCols];
  __shared__ double sh[256];
  if (h_ready)
    for (int j = threadIdx.x; j < k; j += 256) hs[j] = h2[j];
  else
    col_sums(part_in, k, gridDim.x, hs, h2);
  __syncthreads();
  const int e0 = blockIdx.x * R;
  double s2 = 0.0;
  for (int r = threadIdx.x; r < R; r += 256) {
    const int e = e0 + r;
    if (e < n) {
      double v = w[e];
#pragma unroll 8
      for (int j = 0; j < k; j++) v -= hs[j] * Qb[ldq * j + e];
      w[e] = v;
      s2 += v * v;
    }
  }
  s2 = block_sum<256>(s2, sh);
  if (threadIdx.x == 0) norm_part[blockIdx.x] = s2;
}

// The norm of the new Arnoldi vector from its workgroup shares (256
// threads; the same order wherever it is taken).
__device__ double norm_of_parts(const double *norm_part, int nparts, double *sh) {
  double s = 0.0;
  for (int p = threadIdx.x; p < nparts; p += 256) s += norm_part[p];
  return sqrt(block_sum<256>(s, sh));
}

// What the Givens step of Arnoldi iteration `it` reads and writes.
struct GivensArgs {
  const double *h1, *h2;  // the two Gram-Schmidt passes' coefficients
  double *H;
  int ldh;
  double *cs, *sn, *beta, *bpre;
  int *flags;
  double *scal;
  double tol;
  int maxit;
  double switch_rel;
};

// Givens step of Arnoldi iteration `it` (mobility_solver.cpp:1183-1234) on a
// workgroup of 256: H column it-1 = h1 + h2 and the new vector's norm qn,
// rotated by the previous rotations; the new rotation and the residual
// estimate |beta[it]|.  Returns whether the solve has converged (or reached
// maxit).  Only a workgroup with `write` stores the H column, the rotation,
// beta, scal[3] = qn and the flags; the others read nothing that it writes
// (the unrotated right-hand side comes from bpre), so the step may be taken
// redundantly by every workgroup of a kernel.  LDS: hc[kMaxCols + 1],
// scs[kMaxCols], ssn[kMaxCols], one int.
__device__ bool givens_column(double qn, int it, const GivensArgs &g, bool write, double *hc,
                              double *scs, double *ssn, int *sconv) {
  for (int i = threadIdx.x; i < it; i += 256) {
    hc[i] = g.h1[i] + g.h2[i];
    if (i < it - 1) { scs[i] = g.cs[i]; ssn[i] = g.sn[i]; }
  }
  __syncthreads();
  const int c = it - 1;
  if (threadIdx.x == 0) {
    hc[it] = qn;
    for (int i = 1; i < it; i++) {
      const double t = scs[i - 1] * hc[i - 1] + ssn[i - 1] * hc[i];
      hc[i] = scs[i - 1] * hc[i] - ssn[i - 1] * hc[i - 1];
      hc[i - 1] = t;
    }
    double csc, snc;
    if (hc[it] == 0.0) {
      csc = 1.0; snc = 0.0;
    } else if (fabs(hc[it]) > fabs(hc[c])) {
      const double t = hc[c] / hc[it];
      snc = (2.0 * double(hc[it] > 0) - 1.0) / sqrt(1.0 + t * t);
      csc = t * snc;
    } else {
      const double t = hc[it] / hc[c];
      csc = (2.0 * double(hc[c] > 0) - 1.0) / sqrt(1.0 + t * t);
      snc = t * csc;
    }
    hc[c] = csc * hc[c] + snc * hc[it];
    hc[it] = 0.0;
    const double bc = g.bpre[c];
    const double bit = -snc * bc;
    const double rel = fabs(bit) / g.scal[0];
    const bool conv = rel <= g.tol || it == g.maxit;
    *sconv = conv;
    if (write) {
      g.cs[c] = csc; g.sn[c] = snc;
      g.beta[it] = bit;
      g.bpre[it] = bit;
      g.beta[c] = csc * bc;
      g.scal[3] = qn;
      if (conv) {
        g.flags[F_DONE] = 1;
        g.flags[F_ITS] = it;
        g.flags[F_SKIP64] = 1;
        g.flags[F_SKIP32] = 1;
      } else if (rel <= g.switch_rel) {
        // relaxed Krylov: once the residual has dropped below switch_rel the
        // operator may carry a relative error up to tol / rel, so the FP32
        // pair kernel replaces the FP64 one for the remaining Arnoldi steps
        g.flags[F_SKIP64] = 1;
        g.flags[F_SKIP32] = 0;
      }
    }
  }
  __syncthreads();
  if (write)
    for (int i = threadIdx.x; i <= it; i += 256) g.H[size_t(i) * g.ldh + c] = hc[i];
  return *sconv != 0;
}

// Givens step of Arnoldi iteration `it` as a kernel of its own: the new
// vector's norm is the sum of the workgroup shares.  The division by the
// norm happens in the next iteration's first kernel.
__global__ __launch_bounds__(256) void k_givens(const double *norm_part, int nparts, int it,
                                                GivensArgs g) {
  if (g.flags[F_DONE]) return;
  __shared__ double sh[256];
  __shared__ double hc[kMaxCols + 1], scs[kMaxCols], ssn[kMaxCols];
  __shared__ int sconv;
  const double qn = norm_of_parts(norm_part, nparts, sh);
  (void)givens_column(qn, it, g, true, hc, scs, ssn, &sconv);
}

// GMRES start on the device, after k_norms left scal[0] = ||rhs|| and
// scal[1] = ||r0||: a zero right-hand side gives x = 0; a residual within
// the tolerance needs no Arnoldi step (F_DONE); a warm start worse than zero
// or not finite is replaced by zero, r0 = rhs (mobility_solver.cpp:1060-1156
// keeps the previous solution otherwise).  r0 is normalised, its norm goes
// to beta[0] and scal[2], and the flags are reset.
__global__ void k_gmres_start(double *x, double *r, const double *rhs, int n, double *scal,
                              double *beta, double *bpre, int *flags, double tol) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const double nrhs = scal[0], b0 = scal[1];
  const bool zero = nrhs == 0.0;
  const bool conv = zero || fabs(b0) <= tol * nrhs;
  const bool reset = !conv && (!isfinite(b0) || fabs(b0) > nrhs);
  const double bb = reset ? nrhs : b0;
  if (e < n) {
    if (zero || reset) x[e] = 0.0;
    if (!conv) r[e] = (reset ? rhs[e] : r[e]) / bb;
  }
  if (e == 0) {
    scal[2] = bb;
    beta[0] = bb;
    bpre[0] = bb;
    flags[F_DONE] = conv ? 1 : 0;
    flags[F_ITS] = 0;
    flags[F_SKIP64] = 0;
    flags[F_SKIP32] = 1;
  }
}

// back substitution for y (mobility_solver.cpp:1236-1257) by one wave:
// column sweep from the last unknown, each lane updating its rows, which
// subtracts the terms in the reference's order (j from it-1 down to i+1)
__global__ __launch_bounds__(64) void k_backsub(const double *H, int ldh, const double *beta,
                                                const int *flags, double *y, const int *skip) {
  if (skip && *skip) return;
  extern __shared__ double rb[];
  const int it = flags[F_ITS];
  const int lane = threadIdx.x;
  for (int i = lane; i < it; i += 64) rb[i] = beta[i];
  __syncthreads();
  for (int j = it - 1; j >= 0; j--) {
    const double yj = rb[j] / H[size_t(j) * ldh + j];
    for (int i = lane; i < j; i += 64) rb[i] -= H[size_t(i) * ldh + j] * yj;
    if (lane == 0) y[j] = yj;
    __syncthreads();
  }
}

__global__ void k_combine(const double *Qb, size_t ldq, const int *flags, const double *y,
                          double *t, int n, const int *skip) {
  if (skip && *skip) return;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int it = flags[F_ITS];
  double s = 0;
  for (int i = 0; i < it; i++) s += y[i] * Qb[ldq * i + e];
  t[e] = s;
}

// out = a*x + b*y  (elementwise); skip: a no-op if non-null and nonzero
__global__ void k_axpby(double *out, double a, const double *x, double b, const double *y, int n,
                        const int *skip) {
  if (skip && *skip) return;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) out[e] = a * x[e] + (y ? b * y[e] : 0.0);
}

// scal[o0] = ||x||, scal[o1] = ||y|| (one workgroup): four running sums per
// thread over a stride of 4096, then the wave sums and a fixed tree
__global__ __launch_bounds__(1024) void k_norms(const double *x, const double *y, int n,
                                                double *scal, int o0, int o1, const int *skip) {
  if (skip && *skip) return;
  __shared__ double sh[16];
  for (int pass = 0; pass < 2; pass++) {
    const double *v = pass ? y : x;
    if (!v) break;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int e = threadIdx.x;
    for (; e + 3072 < n; e += 4096) {
      const double a = v[e], b = v[e + 1024], c = v[e + 2048], d = v[e + 3072];
      s0 += a * a; s1 += b * b; s2 += c * c; s3 += d * d;
    }
    for (; e < n; e += 1024) s0 += v[e] * v[e];
    const double s = block_sum<1024>((s0 + s1) + (s2 + s3), sh);
    if (threadIdx.x == 0) scal[pass ? o1 : o0] = sqrt(s);
  }
}

__global__ void k_scale_by(double *x, int n, const double *scal, int o) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) x[e] /= scal[o];
}

__global__ void k_set_ints(int *p, int n, int v, const int *skip) {
  if (skip && *skip) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

inline int nblk(int n, int b) { return (n + b - 1) / b; }

// Relative error of the FP32-Kahan pair operator against FP64 (measured
// max-norm error ~6e-7 on 46.7k particles, bounded here by 1e-6).
constexpr double kFp32OperatorError = 1e-6;

static int launch_ok(gmi_ctx *c, const char *what) { return hip_check(c, hipGetLastError(), what); }

static int krylov_read_flags(gmi_ctx *c, KrylovState *s) {
  int rc;
  if ((rc = hip_check(c, hipMemcpyAsync(s->h_flags, s->flags, sizeof(int) * F_N,
                                        hipMemcpyDeviceToHost, c->stream), "flags D2H")))
    return rc;
  return hip_check(c, hipStreamSynchronize(c->stream), "sync");
}

// Row split of the column dot products: nb workgroups of R rows.
// Up to kMaxDotBlocks workgroups every consumer workgroup re-sums a column's
// partials itself; beyond that (wide) a k_colsum_wide launch sums them once
// and up to ~kWideDotBlocks workgroups share the rows.
constexpr int kWideDotBlocks = 4096;
struct DotShape {
  int nb, R;
  bool wide;
};
static DotShape dot_shape(int n) {
  const int chunks = nblk(n, 256);
  if (chunks <= kMaxDotBlocks) return {chunks, 256, false};
  // 256-row chunks per workgroup
  const int per = std::min((chunks + kWideDotBlocks - 1) / kWideDotBlocks, kMaxRows / 256);
  const int R = per * 256;
  return {nblk(n, R), R, true};
}

// GMRES with right preconditioning (mobility_solver.cpp:970-1354) on s->rhs,
// warm-started from s->x; classical Gram-Schmidt Arnoldi (s->gs_passes
// passes per step).  mixed: relaxed precision (FP64
// operator until the residual drops below switch_rel, FP32 after).
// n: vector length (padding rows stay zero); maxit: iteration cap.
// The stages: gmres_begin (initial residual and the start decisions on the
// device), gmres_enqueue (Arnoldi steps it0..it1, every kernel a no-op once
// the device flag says the solve has converged), gmres_chunks (chunks of
// steps from it0, the host reading the flags after each), gmres_finish (the
// least-squares solution and the update of x).

template <class Op>
int gmres_begin(gmi_ctx *c, KrylovState *s, Op &op, int n, double tol) {
  const int nb = nblk(n, 256);
  int rc;
  // r = rhs - A x ; norms ; the start decisions on the device (no host
  // round trip: a solve that needs no step runs one flag-skipped chunk)
  double *r = s->Qk;
  if (op.zero_start()) {
    // an exact preconditioner solves from x0 = 0 in one step: no warm start,
    // so no operator apply for the initial residual (r0 = rhs)
    hipLaunchKernelGGL(k_axpby, dim3(nb), dim3(256), 0, c->stream, s->x, 0.0, s->rhs, 0.0,
                       nullptr, n, nullptr);
    hipLaunchKernelGGL(k_axpby, dim3(nb), dim3(256), 0, c->stream, r, 1.0, s->rhs, 0.0, nullptr,
                       n, nullptr);
  } else {
    if ((rc = op.mult(s->x, s->tmp, 0, nullptr, false))) return rc;
    hipLaunchKernelGGL(k_axpby, dim3(nb), dim3(256), 0, c->stream, r, 1.0, s->rhs, -1.0, s->tmp,
                       n, nullptr);
  }
  hipLaunchKernelGGL(k_norms, dim3(1), dim3(1024), 0, c->stream, s->rhs, r, n, s->scal, 0, 1,
                     nullptr);
  hipLaunchKernelGGL(k_gmres_start, dim3(nb), dim3(256), 0, c->stream, s->x, r, s->rhs, n, s->scal,
                     s->beta, s->bpre, s->flags, tol);
  return launch_ok(c, "gmres start");
}

static GivensArgs givens_args(KrylovState *s, int maxit, bool mixed, double tol) {
  // scal[0] = ||rhs|| is read by k_givens
  return GivensArgs{s->h1, s->h2, s->H, s->cols + 1, s->cs, s->sn, s->beta, s->bpre,
                    s->flags, s->scal, tol, maxit, mixed ? tol / kFp32OperatorError : 0.0};
}

// Arnoldi steps it0..it1 (each kernel a no-op once F_DONE is set)
template <class Op>
int gmres_enqueue(gmi_ctx *c, KrylovState *s, Op &op, int n, int maxit, bool mixed, double tol,
                  int it0, int it1) {
  const size_t ldq = size_t(n);
  const DotShape ds = dot_shape(n);
  const size_t shm = sizeof(double) * ds.R;
  double *pp1 = s->pp, *pp2 = s->pp + s->pp_half;
  const GivensArgs ga = givens_args(s, maxit, mixed, tol);
  int r;
  if (op.fused()) {
    // the operator runs each Arnoldi step on its own row partition (the
    // Givens step of the previous iteration folded into the first kernel)
    // and ends the chunk with a Givens step of its own, so that the host
    // can read the flags
    for (int it = it0; it <= it1; it++)
      if ((r = op.fused_step(s, it, it == it0, mixed, ga))) return r;
    hipLaunchKernelGGL(k_givens, dim3(1), dim3(256), 0, c->stream, s->part, op.fused_parts(),
                       it1, ga);
    return launch_ok(c, "givens");
  }
  for (int it = it0; it <= it1; it++) {
    double *qprev = s->Qk + ldq * (it - 1);
    double *w = s->Qk + ldq * it;
    double *zk = s->Zk ? s->Zk + ldq * (it - 1) : s->z;
    // the previous vector is normalised by the preconditioner's first
    // kernel (the first one was scaled by beta0 already)
    if ((r = op.precond(qprev, zk, s->flags, it >= 2 ? 3 : -1))) return r;
    if ((r = op.mult(zk, w, mixed ? 2 : 0, s->flags, true))) return r;
    // pass-1 dots, update + norm (one pass), or update + pass-2 dots,
    // update + norm (two passes); then Givens
    hipEvent_t gs_stop = nullptr;
    if ((r = prof_gs_start(c, it, n, s->gs_passes, &gs_stop))) return r;
    hipLaunchKernelGGL(k_pdots, dim3(ds.nb), dim3(256), shm, c->stream, s->Qk, ldq, it, w, n,
                       ds.R, pp1, s->flags);
    if (ds.wide)
      hipLaunchKernelGGL(k_colsum_wide, dim3(it), dim3(256), 0, c->stream, pp1, ds.nb, s->h1,
                         s->flags);
    if (s->gs_passes == 1) {
      hipLaunchKernelGGL(k_psub_norm, dim3(ds.nb), dim3(256), 0, c->stream, s->Qk, ldq, it, pp1,
                         s->h1, w, n, ds.R, s->part, s->flags, int(ds.wide), s->h2);
    } else {
      hipLaunchKernelGGL(k_psub_dots, dim3(ds.nb), dim3(256), shm, c->stream, s->Qk, ldq, it,
                         pp1, s->h1, w, n, ds.R, pp2, s->flags, int(ds.wide));
      if (ds.wide)
        hipLaunchKernelGGL(k_colsum_wide, dim3(it), dim3(256), 0, c->stream, pp2, ds.nb, s->h2,
                           s->flags);
      hipLaunchKernelGGL(k_psub_norm, dim3(ds.nb), dim3(256), 0, c->stream, s->Qk, ldq, it, pp2,
                         s->h2, w, n, ds.R, s->part, s->flags, int(ds.wide), nullptr);
    }
    if (gs_stop && (r = hip_check(c, hipEventRecord(gs_stop, c->stream), "event"))) return r;
    hipLaunchKernelGGL(k_givens, dim3(1), dim3(256), 0, c->stream, s->part, ds.nb, it, ga);
    if ((r = launch_ok(c, "arnoldi"))) return r;
  }
  return GMI_OK;
}

// Chunks of Arnoldi steps from it0 until the device flags say the solve has
// converged; *its = the solve's iteration count.
template <class Op>
int gmres_chunks(gmi_ctx *c, KrylovState *s, Op &op, int n, int maxit, bool mixed, double tol,
                 int it0, int *its) {
  int rc;
  if (s->graph_gen != c->list_generation) {
    s->drop_graphs();
    s->graph_gen = c->list_generation;
  }
  // Gram-Schmidt records of this solve start here; steps launched past
  // convergence (flag-skipped) are dropped once the count is known
  const size_t gs_mark = c->gs_recs.size();
  int it = it0;
  int done = 0;
  while (!done) {
    const int last = std::min(maxit, it + s->chunk - 1);
    if (s->use_graphs) {
      // one captured graph per (precision mode, chunk start); replayed by
      // every later solve of this context
      const std::pair<int, int> key{mixed ? 1 : 0, it};
      auto g = s->graphs.find(key);
      if (g == s->graphs.end()) {
        CapturedGraph cg;
        if ((rc = capture_graph(
                 c, [&]() { return gmres_enqueue(c, s, op, n, maxit, mixed, tol, it, last); },
                 &cg)))
          return rc;
        g = s->graphs.emplace(key, cg).first;
      }
      if ((rc = hip_check(c, hipGraphLaunch(g->second.exec, c->stream), "graph launch")))
        return rc;
    } else {
      if ((rc = gmres_enqueue(c, s, op, n, maxit, mixed, tol, it, last))) return rc;
    }
    it = last + 1;
    if ((rc = krylov_read_flags(c, s))) return rc;
    done = s->h_flags[F_DONE];
    if (!done && it > maxit) return fail(c, GMI_ERR_STATE, "GMRES ran past its iteration cap");
  }
  const int nits = s->h_flags[F_ITS];
  c->gs_recs.erase(std::remove_if(c->gs_recs.begin() + gs_mark, c->gs_recs.end(),
                                  [nits](const gmi_ctx::GsRec &g) { return g.k > nits; }),
                   c->gs_recs.end());
  *its = nits;
  return GMI_OK;
}

// The least-squares solution and x += P^-1 Q y (or x += Z y); skip gates
// every kernel (null: always run).
template <class Op>
int gmres_finish(gmi_ctx *c, KrylovState *s, Op &op, int n, const int *skip) {
  const size_t ldq = size_t(n);
  const int nb = nblk(n, 256);
  const int ldh = s->cols + 1;
  int rc;
  hipLaunchKernelGGL(k_backsub, dim3(1), dim3(64), sizeof(double) * (s->cols + 1), c->stream,
                     s->H, ldh, s->beta, s->flags, s->y, skip);
  if (s->Zk) {
    // x += Z y
    hipLaunchKernelGGL(k_combine, dim3(nb), dim3(256), 0, c->stream, s->Zk, ldq, s->flags, s->y,
                       s->z, n, skip);
    if ((rc = launch_ok(c, "gmres solution"))) return rc;
  } else {
    // x += P^-1 (Q y)
    hipLaunchKernelGGL(k_combine, dim3(nb), dim3(256), 0, c->stream, s->Qk, ldq, s->flags, s->y,
                       s->tmp, n, skip);
    if ((rc = launch_ok(c, "gmres solution"))) return rc;
    if ((rc = op.precond(s->tmp, s->z, skip, -1))) return rc;
  }
  hipLaunchKernelGGL(k_axpby, dim3(nb), dim3(256), 0, c->stream, s->x, 1.0, s->x, 1.0, s->z, n,
                     skip);
  return launch_ok(c, "gmres update");
}

template <class Op>
int krylov_gmres(gmi_ctx *c, KrylovState *s, Op &op, int n, int maxit, bool mixed, double tol,
                 int *its) {
  int rc;
  *its = 0;
  if ((rc = gmres_begin(c, s, op, n, tol))) return rc;
  int nits = 0;
  if ((rc = gmres_chunks(c, s, op, n, maxit, mixed, tol, 1, &nits))) return rc;
  if ((rc = gmres_finish(c, s, op, n, nullptr))) return rc;
  *its = nits;
  return GMI_OK;
}

// Linear solve.  FP64: plain GMRES.  Mixed: relaxed-precision GMRES, then
// the true residual with the FP64 operator; if it misses the tolerance GMRES
// restarts from the current iterate (iterative refinement, FP64 outer).
template <class Op>
int krylov_solve(gmi_ctx *c, KrylovState *s, Op &op, int n, int maxit, int precision, double tol,
                 int *its) {
  // an operator whose preconditioner is (nearly) exact keeps its Krylov
  // products in FP64: GMRES then takes one or two steps, and the refinement
  // pass would cost more than the FP32 products save
  if (precision == GMI_FP64 || op.fp64_krylov())
    return krylov_gmres(c, s, op, n, maxit, false, tol, its);
  const int nb = nblk(n, 256);
  int rc, total = 0;
  for (int cycle = 0; cycle < 4; cycle++) {
    int it = 0;
    if ((rc = krylov_gmres(c, s, op, n, maxit, true, tol, &it))) return rc;
    total += it;
    // true residual with the FP64 operator
    if ((rc = op.mult(s->x, s->tmp, 0, nullptr, false))) return rc;
    hipLaunchKernelGGL(k_axpby, dim3(nb), dim3(256), 0, c->stream, s->z, 1.0, s->rhs, -1.0,
                       s->tmp, n, nullptr);
    hipLaunchKernelGGL(k_norms, dim3(1), dim3(1024), 0, c->stream, s->rhs, s->z, n, s->scal, 4, 5,
                       nullptr);
    if ((rc = launch_ok(c, "refinement residual"))) return rc;
    if ((rc = hip_check(c, hipMemcpyAsync(s->h_scal + 4, s->scal + 4, 2 * sizeof(double),
                                          hipMemcpyDeviceToHost, c->stream), "scal")))
      return rc;
    if ((rc = hip_check(c, hipStreamSynchronize(c->stream), "sync"))) return rc;
    const double nb0 = s->h_scal[4], nr = s->h_scal[5];
    if (nb0 == 0.0 || nr <= tol * nb0) break;
  }
  *its = total;
  return GMI_OK;
}

}  // namespace
}  // namespace gmi
