// gmi_seed.hip -- seeding points on the unit sphere as the reference does
// (seeding.cu:13-389, cuda_functions.cu:1159-1250): the uniform seeding runs
// the reference's repulsion relaxation on the GPU; the equatorial and
// "platy" seedings are closed-form spirals computed on the host.
//
// Uniform (repulsive) seeding, per iteration (Jacobi style, as the
// reference's GPU path):
//   F_i = sum_{j != i} 5 (2 - |d|) d/|d|,  d = x_i - x_j   (all pairs; on the
//         unit sphere every |d| <= 2, so every pair repels)
//   x_i <- (x_i + F_i) / |x_i + F_i|
// until the largest change of any coordinate is at most 1e-4.  The start is
// the reference's spiral: z_n = 2n/(N-1) - 1, phi advancing by
// 3.6/(r sqrt(N)) (seeding.cu:19-38).
#include "gmi_ctx.hpp"
#include "gmi_sim.h"

#include <cmath>
#include <vector>

namespace gmi {
namespace {

constexpr int kSeedBlock = 256;
constexpr double kPiSeed = 3.14159265358979323846;

__global__ __launch_bounds__(kSeedBlock) void k_seed_force(const double *__restrict__ X, int N,
                                                           double *__restrict__ F) {
  __shared__ double sx[kSeedBlock], sy[kSeedBlock], sz[kSeedBlock];
  const int i = blockIdx.x * kSeedBlock + threadIdx.x;
  const int ii = i < N ? i : 0;
  const double xi = X[3 * ii], yi = X[3 * ii + 1], zi = X[3 * ii + 2];
  double fx = 0.0, fy = 0.0, fz = 0.0;
  for (int base = 0; base < N; base += kSeedBlock) {
    const int n = min(kSeedBlock, N - base);
    __syncthreads();
    if (threadIdx.x < n) {
      const int j = base + threadIdx.x;
      sx[threadIdx.x] = X[3 * j]; sy[threadIdx.x] = X[3 * j + 1]; sz[threadIdx.x] = X[3 * j + 2];
    }
    __syncthreads();
    for (int jj = 0; jj < n; jj++) {
      if (base + jj == i) continue;
      const double dx = xi - sx[jj], dy = yi - sy[jj], dz = zi - sz[jj];
      const double dist = sqrt(dx * dx + dy * dy + dz * dz);
      const double dm1 = 1.0 / dist;
      const double fac = 5.0 * (2.0 - dist);
      fx += fac * dx * dm1;
      fy += fac * dy * dm1;
      fz += fac * dz * dm1;
    }
  }
  if (i < N) { F[3 * i] = fx; F[3 * i + 1] = fy; F[3 * i + 2] = fz; }
}

// x <- normalised (x + F); per workgroup the largest coordinate change
__global__ __launch_bounds__(kSeedBlock) void k_seed_move(double *__restrict__ X,
                                                          const double *__restrict__ F, int N,
                                                          double *__restrict__ bmax) {
  __shared__ double sh[kSeedBlock];
  const int i = blockIdx.x * kSeedBlock + threadIdx.x;
  double mv = 0.0;
  if (i < N) {
    double p[3] = {X[3 * i] + F[3 * i], X[3 * i + 1] + F[3 * i + 1], X[3 * i + 2] + F[3 * i + 2]};
    const double nrm = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    for (int k = 0; k < 3; k++) {
      p[k] /= nrm;
      mv = fmax(mv, fabs(p[k] - X[3 * i + k]));
      X[3 * i + k] = p[k];
    }
  }
  sh[threadIdx.x] = mv;
  __syncthreads();
  for (int s = kSeedBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) bmax[blockIdx.x] = sh[0];
}

// seeding.cu:19-38 (and :271-290 for filaments): the starting spiral
void spiral_start(int N, double *out) {
  double phi = 0.0;
  for (int n = 0; n < N; n++) {
    const double z = (N == 1) ? -1.0 : 2.0 * n / double(N - 1) - 1.0;
    const double r = std::sqrt(1.0 - z * z);
    if (n == 0 || n == N - 1)
      phi = 0.0;
    else
      phi += 3.6 / (r * std::sqrt(double(N)));
    out[3 * n] = r * std::cos(phi);
    out[3 * n + 1] = r * std::sin(phi);
    out[3 * n + 2] = z;
  }
}

// one band of the equatorial spiral (seeding.cu:147-172): n points from
// `first` on a spiral laid out for `total` points over the whole sphere
void band_spiral(int n, int total, int offset, double *out) {
  double phi = 0.0;
  for (int k = 0; k < n; k++) {
    const double z = 2.0 * (k + offset) / double(total - 1) - 1.0;
    const double r = std::sqrt(1.0 - z * z);
    out[3 * k] = r * std::cos(phi);
    out[3 * k + 1] = r * std::sin(phi);
    out[3 * k + 2] = z;
    phi += 3.6 / (r * std::sqrt(double(total)));
  }
}

}  // namespace
}  // namespace gmi

using namespace gmi;

extern "C" {

int gmi_seed_sphere(gmi_ctx *c, int n, int kind, double r_over_l, int max_iterations,
                    double *unit_xyz, int *iterations) {
  if (!c) return GMI_ERR_ARG;
  if (n < 1 || !unit_xyz || max_iterations < 1) return fail(c, GMI_ERR_ARG, "bad seeding arguments");
  if (iterations) *iterations = 0;
  if (kind == GMI_SEED_EQUATORIAL || kind == GMI_SEED_PLATY) {
    if (!(r_over_l > 0)) return fail(c, GMI_ERR_ARG, "R_OVER_L must be positive");
    const double theta_max = 0.5 / r_over_l;  // band width L across the surface
    const double h = 2.0 * std::sin(theta_max);
    const int in_band = kind == GMI_SEED_PLATY ? int(std::lround(0.9 * n)) : n;
    const int total = int(std::ceil(in_band * 2.0 / h));
    if (total < 2 || total < in_band) return fail(c, GMI_ERR_ARG, "band wider than the sphere");
    const int offset = int(std::lround(0.5 * (total - in_band)));
    band_spiral(in_band, total, offset, unit_xyz);
    // platy: the remaining tenth on a ring at z = -0.9 (seeding.cu:214-226)
    for (int k = in_band; k < n; k++) {
      const double z = -0.9, r = std::sqrt(1.0 - z * z);
      const double phi = 2.0 * kPiSeed * double(k - in_band) / double(n - in_band);
      unit_xyz[3 * k] = r * std::cos(phi);
      unit_xyz[3 * k + 1] = r * std::sin(phi);
      unit_xyz[3 * k + 2] = z;
    }
    return GMI_OK;
  }
  if (kind != GMI_SEED_UNIFORM) return fail(c, GMI_ERR_ARG, "unknown seeding kind");
  spiral_start(n, unit_xyz);
  int rc;
  if ((rc = hip_check(c, hipSetDevice(c->device), "set device"))) return rc;
  const int nb = (n + kSeedBlock - 1) / kSeedBlock;
  double *X = nullptr, *F = nullptr, *bmax = nullptr, *hmax = nullptr;
  auto cleanup = [&]() {
    if (X) (void)hipFree(X);
    if (F) (void)hipFree(F);
    if (bmax) (void)hipFree(bmax);
    if (hmax) (void)hipHostFree(hmax);
  };
  if ((rc = hip_check(c, hipMalloc(&X, sizeof(double) * 3 * n), "hipMalloc seed")) ||
      (rc = hip_check(c, hipMalloc(&F, sizeof(double) * 3 * n), "hipMalloc seed")) ||
      (rc = hip_check(c, hipMalloc(&bmax, sizeof(double) * nb), "hipMalloc seed")) ||
      (rc = hip_check(c, hipHostMalloc((void **)&hmax, sizeof(double) * nb, hipHostMallocDefault),
                      "pinned seed"))) {
    cleanup();
    return rc;
  }
  if ((rc = hip_check(c, hipMemcpyAsync(X, unit_xyz, sizeof(double) * 3 * n, hipMemcpyHostToDevice,
                                        c->stream), "seed H2D"))) {
    cleanup();
    return rc;
  }
  double biggest = 1e4;
  int it = 0;
  while (biggest > 1e-4 && it < max_iterations) {
    hipLaunchKernelGGL(k_seed_force, dim3(nb), dim3(kSeedBlock), 0, c->stream, X, n, F);
    hipLaunchKernelGGL(k_seed_move, dim3(nb), dim3(kSeedBlock), 0, c->stream, X, F, n, bmax);
    if ((rc = hip_check(c, hipGetLastError(), "seed launch")) ||
        (rc = hip_check(c, hipMemcpyAsync(hmax, bmax, sizeof(double) * nb, hipMemcpyDeviceToHost,
                                          c->stream), "seed D2H")) ||
        (rc = hip_check(c, hipStreamSynchronize(c->stream), "seed sync"))) {
      cleanup();
      return rc;
    }
    biggest = 0.0;
    for (int b = 0; b < nb; b++) biggest = std::fmax(biggest, hmax[b]);
    it++;
  }
  rc = hip_check(c, hipStreamSynchronize(c->stream), "seed sync");
  if (!rc)
    rc = hip_check(c, hipMemcpy(unit_xyz, X, sizeof(double) * 3 * n, hipMemcpyDeviceToHost),
                   "seed D2H");
  cleanup();
  if (rc) return rc;
  if (iterations) *iterations = it;
  if (biggest > 1e-4) return fail(c, GMI_ERR_NOT_CONVERGED, "seeding did not converge");
  return GMI_OK;
}

}  // extern "C"
