// gmi_gemm.hip -- FP64 matrix products on the gfx950 matrix cores.
//
// C = alpha op(A) B + beta C, column-major, op(A) = A or A^T.  The one-time
// set-up of the body-frame preconditioner (gmi_dense.hip) inverts each
// swimmer's blob mobility by block Gauss-Jordan; its rank-128 updates and
// the few skinny products around them run here, in place of the LAPACK /
// BLAS routines the reference calls (matrix.cpp:538-641).
//
// A workgroup of four waves owns a 64 x 64 tile of C; wave w a 32 x 32
// quarter of it, as 2 x 2 blocks of 16 x 16 held in the accumulators of
// v_mfma_f64_16x16x4f64.  The k dimension runs in steps of 16: the
// workgroup stages op(A)[tile rows, k..k+15] and B[k..k+15, tile cols] in
// LDS (coalesced global reads), then each wave issues 4 x 4 matrix
// instructions.  Operand layout of the 16x16x4 FP64 instruction (checked
// on the device with one-hot operands): lane l holds A(l % 16, l / 16) and
// B(l / 16, l % 16); the accumulator holds D(l / 16 + 4 v, l % 16) in
// element v.
#include "gmi_internal.hpp"

namespace gmi {
namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int kGT = 64;   // C tile edge
constexpr int kGK = 16;   // k step staged in LDS

template <bool TA>
__global__ __launch_bounds__(256) void k_dgemm(int m, int n, int k, double alpha,
                                               const double *__restrict__ A, size_t lda,
                                               const double *__restrict__ B, size_t ldb,
                                               double beta, double *__restrict__ C, size_t ldc) {
  __shared__ double sA[kGK][kGT + 1];  // op(A)(i0 + i, k0 + kk) at [kk][i]
  __shared__ double sB[kGK][kGT + 1];  // B(k0 + kk, j0 + j) at [kk][j]
  const int i0 = blockIdx.x * kGT, j0 = blockIdx.y * kGT;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wi = (wave & 1) * 32, wj = (wave >> 1) * 32;
  const int r = lane & 15, q = lane >> 4;
  v4d acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) acc[a][b] = v4d{0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < k; k0 += kGK) {
#pragma unroll
    for (int e = t; e < kGK * kGT; e += 256) {
      int ii, kk;
      if (TA) { kk = e & (kGK - 1); ii = e / kGK; }   // A^T: k runs along memory
      else { ii = e & (kGT - 1); kk = e / kGT; }       // A: i runs along memory
      const int gi = i0 + ii, gk = k0 + kk;
      double v = 0.0;
      if (gi < m && gk < k) v = TA ? A[gk + gi * lda] : A[gi + gk * lda];
      sA[kk][ii] = v;
    }
#pragma unroll
    for (int e = t; e < kGK * kGT; e += 256) {
      const int kk = e & (kGK - 1), jj = e / kGK;
      const int gk = k0 + kk, gj = j0 + jj;
      sB[kk][jj] = (gk < k && gj < n) ? B[gk + gj * ldb] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < kGK; ks += 4) {
      const double a0 = sA[ks + q][wi + r], a1 = sA[ks + q][wi + 16 + r];
      const double b0 = sB[ks + q][wj + r], b1 = sB[ks + q][wj + 16 + r];
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int gi = i0 + wi + 16 * a + q + 4 * v, gj = j0 + wj + 16 * b + r;
        if (gi < m && gj < n) {
          double *c = C + gi + gj * ldc;
          *c = beta == 0.0 ? alpha * acc[a][b][v] : alpha * acc[a][b][v] + beta * *c;
        }
      }
}

}  // namespace

hipError_t launch_dgemm(bool transA, int m, int n, int k, double alpha, const double *A,
                        size_t lda, const double *B, size_t ldb, double beta, double *C,
                        size_t ldc, hipStream_t s) {
  if (m <= 0 || n <= 0) return hipSuccess;
  const dim3 grid((m + kGT - 1) / kGT, (n + kGT - 1) / kGT);
  if (grid.y > 65535) return hipErrorInvalidValue;
  if (transA)
    hipLaunchKernelGGL(k_dgemm<true>, grid, dim3(256), 0, s, m, n, k, alpha, A, lda, B, ldb,
                       beta, C, ldc);
  else
    hipLaunchKernelGGL(k_dgemm<false>, grid, dim3(256), 0, s, m, n, k, alpha, A, lda, B, ldb,
                       beta, C, ldc);
  return hipGetLastError();
}

}  // namespace gmi
