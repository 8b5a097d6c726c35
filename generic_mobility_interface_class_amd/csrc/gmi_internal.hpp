// gmi_internal.hpp -- shared declarations of the gfx950 library (not public).
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>
#include <map>

namespace gmi {

// Targets per pair-kernel workgroup (one target per lane, 4 waves).
constexpr int kBlock = 256;
// Sources staged in LDS per tile.
constexpr int kTile = 256;

enum PairMask : int { kSS = 1, kSB = 2, kBS = 4, kBB = 8 };

// One target tile: up to kBlock consecutive targets of one species inside one
// swimmer group, and the source ranges of that group.
struct TileDesc {
  int species;        // 0: segment targets, 1: blob targets
  int t0, nt;         // target range in the species array
  int sseg0, nsseg;   // segment sources of the group
  int sblob0, nsblob; // blob sources of the group
  int nsplit;         // source splits (>1: partials reduced afterwards)
  int part0;          // first partial slot of this tile
};

struct WorkItem {
  int tile;
  int split;
};

// Device-side SoA views.  Component c of particle i lives at p[c*ld + i].
struct PairArgs {
  const double *Xs; const double *Fs; int lds;
  const double *Xb; const double *Fb; int ldb;
  double *Vs; double *Vb;
  double *part;             // [slot][6][kBlock] partial sums
  const TileDesc *tiles;
  const WorkItem *work;
  int mask;                 // PairMask bits
  int accumulate;           // 0: V = M F, 1: V += M F
  double a_s, a_b, mu;
  const int *skip = nullptr; // if non-null and *skip != 0 the launch is a no-op
  int tt_only = 0;           // 1: translations only -- segments carry forces
                             // (3 rows of Fs) and get velocities (3 rows of Vs)
  const int *sel32 = nullptr; // launch_pair_dyn: FP32-Kahan iff *sel32 == 0
};

// Work list of one pair launch (built once per mask).
struct WorkList {
  TileDesc *d_tiles = nullptr;
  WorkItem *d_work = nullptr;
  int ntiles = 0, nwork = 0, nslots = 0, nsplit_tiles = 0;
  std::vector<TileDesc> h_tiles;
};

// Launchers (gmi_pair.hip).  Split partials are summed in split order by
// k_reduce over the first ntiles_reduce tiles (0: the caller sums them).
// ev_mid (optional) is recorded between the pair kernel and the reduction.
hipError_t launch_pair(const PairArgs &a, int nwork, int ntiles_reduce, bool fp32,
                       hipStream_t s, hipEvent_t ev_mid = nullptr);
// One launch whose precision a device flag picks (a.sel32); split partials
// summed as by launch_pair.
hipError_t launch_pair_dyn(const PairArgs &a, int nwork, int ntiles_reduce, hipStream_t s);
// Segment-segment mobility above a no-slip wall at z = 0 (FP64; kSS lists).
hipError_t launch_pair_wall(const PairArgs &a, int nwork, int ntiles_reduce, hipStream_t s,
                            hipEvent_t ev_mid = nullptr);
// Hashed cell grid of the steric barrier: H buckets (a power of two and a
// multiple of kScanChunk), cells of side 1/inv_h, per-particle group (the
// swimmer when swimmers are independent, else 0).  cnt and fill must be zero
// before a build; every build leaves them zero.
constexpr int kScanChunk = 4096;
struct CellGrid {
  int *cnt, *fill, *start, *boff, *bucket;
  int *arr;         // per bucket entry: particle index in arrival order
  int *ent;         // per bucket entry: particle index in index order
  int4 *key;        // per particle: cell x, y, z and group
  // per bucket entry, in bucket order: position, cell and group
  double *rx, *ry, *rz;
  int4 *rk;
  const int *grp;
  int H;
  double inv_h;
};
// Barrier forces of every particle (Ns segments, then Nb blobs): segments add
// theirs to Fs, blobs write theirs to Frep.
hipError_t launch_barrier(double *Fs, int lds, double *Frep_b, int ldb, const double *Xs,
                          const double *Xb, int Ns, int Nb, const CellGrid &G,
                          int nseg_per_fil, double a_s, double a_b, double K, hipStream_t s);

// C = alpha op(A) B + beta C, column-major FP64 on the matrix cores
// (gmi_gemm.hip); op(A) = A (m x k) or, with transA, A^T (A is k x m).
hipError_t launch_dgemm(bool transA, int m, int n, int k, double alpha, const double *A,
                        size_t lda, const double *B, size_t ldb, double beta, double *C,
                        size_t ldc, hipStream_t s);

// AoS <-> SoA transposes (gmi_pair.hip).
hipError_t launch_aos_to_soa(double *dst, int ld, const double *src, int n,
                             int ncomp, hipStream_t s);
hipError_t launch_soa_to_aos(double *dst, const double *src, int ld, int n,
                             int ncomp, hipStream_t s);

}  // namespace gmi
