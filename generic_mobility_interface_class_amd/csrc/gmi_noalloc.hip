// gmi_noalloc.hip -- the grand-mobility apply for callers that cannot
// allocate (code under HIP graph capture, serving loops, harnesses that
// forbid hipMalloc).
//
// gmi_evaluate_grand_mobility_noalloc runs the same pair kernels as
// gmi_evaluate_grand_mobility (k_pair and its split reduction), on device
// arrays the caller owns.  The only device memory beyond the caller's inputs
// and outputs -- the target-tile table, the work list and the split
// partials -- lives in `scratch`, a file-scope device array of fixed size.
// The host never names `scratch` itself: each device's address of it comes
// from hipGetSymbolAddress, taken once per device (with that device current)
// on the first call there, and every kernel receives that address.
//
// Every call on one device shares that device's scratch, which keeps the
// contents of the last call.  Nothing in it has to start at zero: a call
// writes the tile table, the work list and every partial slot it reads
// before reading it.  Calls that overlap on one device (other streams, other
// host threads) race on the scratch; calls on one stream are ordered.
#include "gmi_internal.hpp"
#include "../../include/gmi.h"

#include <mutex>

namespace gmi {
namespace {

// Launch shape of the context's default work lists (gmi_ctx: 2048 target
// workgroups, source chunks of at least 128), so that a call splits its
// sources exactly as gmi_evaluate_grand_mobility does on a context with the
// default shape.
constexpr int kNaTargetBlocks = 2048;
constexpr int kNaMinChunk = 128;

constexpr int kNaMaxTiles = GMI_NOALLOC_MAX_TILES;   // target tiles of kBlock
constexpr int kNaMaxSlots = 2560;                    // split partial slots
constexpr size_t kNaPartBytes = size_t(kNaMaxSlots) * 6 * kBlock * sizeof(double);
constexpr size_t kNaTileBytes = size_t(kNaMaxTiles) * sizeof(TileDesc);
// nsplit > 1 only while tiles * nsplit <= kNaMaxSlots, else one split per tile
constexpr int kNaMaxWork = kNaMaxSlots > kNaMaxTiles ? kNaMaxSlots : kNaMaxTiles;
constexpr size_t kNaWorkBytes = size_t(kNaMaxWork) * sizeof(WorkItem);
constexpr size_t kNaTileOff = kNaPartBytes;
constexpr size_t kNaWorkOff = (kNaTileOff + kNaTileBytes + 255) / 256 * 256;
constexpr size_t kNaBytes = kNaWorkOff + kNaWorkBytes;
static_assert(kNaPartBytes % 256 == 0, "partials end on a 256-byte boundary");
static_assert(kNaBytes == GMI_NOALLOC_SCRATCH_BYTES, "scratch size in gmi.h");

}  // namespace
}  // namespace gmi

// 31,637,504 bytes per device (30.2 MiB), held until the process ends.
static __device__ __align__(256) unsigned char scratch[gmi::kNaBytes];

namespace gmi {
namespace {

// The work list of one all-blocks apply over a single group of particles:
// segment tiles first, then blob tiles, nsplit source splits per tile.
__global__ void k_noalloc_worklist(TileDesc *tiles, WorkItem *work, int ns, int nb,
                                   int nst, int ntiles, int nsplit) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= ntiles * nsplit) return;
  const int t = w / nsplit, s = w - t * nsplit;
  work[w] = {t, s};
  if (s != 0) return;
  const bool seg = t < nst;
  const int first = (seg ? t : t - nst) * kBlock;
  TileDesc T;
  T.species = seg ? 0 : 1;
  T.t0 = first;
  T.nt = min(kBlock, (seg ? ns : nb) - first);
  T.sseg0 = 0; T.nsseg = ns;
  T.sblob0 = 0; T.nsblob = nb;
  T.nsplit = nsplit;
  T.part0 = nsplit > 1 ? t * nsplit : 0;
  tiles[t] = T;
}

std::mutex g_scratch_mu;
void *g_scratch[GMI_NOALLOC_MAX_DEVICES];

// This device's address of `scratch`, taken on the first call with the
// device current.
int scratch_of_current_device(unsigned char **out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    (void)hipGetLastError();
    return GMI_ERR_HIP;
  }
  if (dev < 0 || dev >= GMI_NOALLOC_MAX_DEVICES) return GMI_ERR_ARG;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return GMI_ERR_HIP;
    }
    g_scratch[dev] = p;
  }
  *out = static_cast<unsigned char *>(g_scratch[dev]);
  return GMI_OK;
}

}  // namespace
}  // namespace gmi

extern "C" int gmi_evaluate_grand_mobility_noalloc(
    const double *x_segs, const double *f_segs, int n_segs, int ld_segs,
    const double *x_blobs, const double *f_blobs, int n_blobs, int ld_blobs,
    double *v_segs, double *v_blobs, double mu, double a_seg, double a_blob, int precision,
    void *stream) {
  using namespace gmi;
  if (n_segs < 0 || n_blobs < 0 || ld_segs < n_segs || ld_blobs < n_blobs ||
      !(mu > 0) || !(a_seg > 0) || !(a_blob > 0) ||
      (precision != GMI_FP64 && precision != GMI_FP32_KAHAN))
    return GMI_ERR_ARG;
  if (n_segs > 0 && (!x_segs || !f_segs || !v_segs)) return GMI_ERR_ARG;
  if (n_blobs > 0 && (!x_blobs || !f_blobs || !v_blobs)) return GMI_ERR_ARG;
  const long long nst = (n_segs + kBlock - 1) / kBlock;
  const long long ntiles = nst + (n_blobs + kBlock - 1) / kBlock;
  if (ntiles > kNaMaxTiles) return GMI_ERR_TOO_LARGE;
  if (ntiles == 0) return GMI_OK;
  // source splits as the context's work list builds them (one group)
  const long long L = (long long)n_segs + n_blobs;
  const long long base = std::max<long long>(1, (kNaTargetBlocks + ntiles - 1) / ntiles);
  const long long most = std::max<long long>(1, (L + kNaMinChunk - 1) / kNaMinChunk);
  int nsplit = int(std::max<long long>(1, std::min(base, most)));
  if (nsplit > 1 && ntiles * nsplit > kNaMaxSlots) nsplit = int(std::max<long long>(1, kNaMaxSlots / ntiles));
  const int nwork = int(ntiles) * nsplit;

  unsigned char *sp = nullptr;
  int rc;
  if ((rc = scratch_of_current_device(&sp))) return rc;
  double *part = reinterpret_cast<double *>(sp);
  TileDesc *tiles = reinterpret_cast<TileDesc *>(sp + kNaTileOff);
  WorkItem *work = reinterpret_cast<WorkItem *>(sp + kNaWorkOff);
  hipStream_t s = static_cast<hipStream_t>(stream);

  hipLaunchKernelGGL(k_noalloc_worklist, dim3((nwork + 255) / 256), dim3(256), 0, s, tiles, work,
                     n_segs, n_blobs, int(nst), int(ntiles), nsplit);
  if (hipGetLastError() != hipSuccess) return GMI_ERR_HIP;
  PairArgs a;
  a.Xs = x_segs; a.Fs = f_segs; a.lds = std::max(ld_segs, 1);
  a.Xb = x_blobs; a.Fb = f_blobs; a.ldb = std::max(ld_blobs, 1);
  a.Vs = v_segs; a.Vb = v_blobs;
  a.part = part;
  a.tiles = tiles;
  a.work = work;
  a.mask = kSS | kSB | kBS | kBB;
  a.accumulate = 0;
  a.a_s = a_seg; a.a_b = a_blob; a.mu = mu;
  if (launch_pair(a, nwork, nsplit > 1 ? int(ntiles) : 0, precision == GMI_FP32_KAHAN, s) !=
      hipSuccess) {
    (void)hipGetLastError();
    return GMI_ERR_HIP;
  }
  return GMI_OK;
}
