// gmi_ctx.hpp -- the opaque context behind the C ABI (not public).
#pragma once
#include "gmi_internal.hpp"

namespace gmi {
// what gmi_ctx::extra holds
enum ExtraKind : int { kExtraNone = 0, kExtraSim = 1, kExtraCilia = 2 };
}  // namespace gmi

struct gmi_ctx {
  // pinned/device staging slots for host copies
  enum { kStXs, kStFs, kStXb, kStFb, kStRep, kStVb, kStVs, kNumStage };
  struct Stage {
    double *pinned = nullptr;
    double *dev = nullptr;
    hipEvent_t done = nullptr;
  };
  struct Pending {
    void *dst;
    const void *src;
    size_t bytes;
  };

  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;

  // sizes
  int nsw = 0;
  std::vector<int> nseg, nblob, seg_off, blob_off;
  int Ns = 0, Nb = 0, lds = 64, ldb = 64;
  int nseg_fil = 1;
  bool interact = true;
  double mu = 1, as = 1, ab = 1;
  int precision = 0;

  // device SoA vectors
  double *Xs = nullptr, *Fs = nullptr, *Vs = nullptr;
  double *Xb = nullptr, *Fb = nullptr, *Vb = nullptr, *Frep = nullptr;
  double *part = nullptr;
  size_t part_slots = 0;

  // launch shaping (set by the tuner)
  bool wall = false;  // segment-segment mobility above a no-slip wall at z = 0
  bool tt_only = false;  // translation-only pair blocks (prescribed cilia)
  int target_blocks = 2048;
  int min_chunk = 128;
  int list_generation = 0;  // bumped whenever the work lists are rebuilt

  std::map<int, gmi::WorkList> lists;
  gmi::CellGrid cells{};  // the barrier's cell grid (device buffers)
  Stage stage[kNumStage];
  std::vector<Pending> pending;
  std::string err;

  // pair-kernel profiling: events around every pair-kernel launch
  bool profiling = false;
  std::vector<hipEvent_t> prof_events;  // pairs (start, stop)
  size_t prof_used = 0;
  // Gram-Schmidt brackets of eager Arnoldi steps while profiling: events
  // gs_events[2i], gs_events[2i+1] around the step's Gram-Schmidt kernels,
  // the step's column count k, vector length n and Gram-Schmidt passes
  struct GsRec {
    size_t ev;
    int k, n, passes;
  };
  std::vector<hipEvent_t> gs_events;
  size_t gs_used = 0;
  std::vector<GsRec> gs_recs;

  // hook for subsystems that own further device memory
  void (*extra_release)(gmi_ctx *) = nullptr;
  void *extra = nullptr;
  int extra_kind = gmi::kExtraNone;
  void release_extra() {
    if (extra_release) extra_release(this);
    extra_release = nullptr;
    extra = nullptr;
    extra_kind = gmi::kExtraNone;
  }
};

namespace gmi {
int fail(gmi_ctx *c, int code, const std::string &msg);
int hip_check(gmi_ctx *c, hipError_t e, const char *what);
int rebuild_lists(gmi_ctx *c);
// Pair-kernel profiling (gmi_profile_begin): records the start event on the
// context stream and returns in *ev_stop the event to record right after the
// pair kernel; nullptr when not profiling or while the stream is captured.
int prof_start(gmi_ctx *c, hipEvent_t *ev_stop);
// Gram-Schmidt profiling of one Arnoldi step with k previous columns of
// length n: records the start event and returns the stop event in *ev_stop
// (nullptr when not profiling or while the stream is captured).
int prof_gs_start(gmi_ctx *c, int k, int n, int passes, hipEvent_t *ev_stop);
// skip: the launch is a no-op when non-null and *skip != 0
int run_pairs(gmi_ctx *c, int mask, int accumulate, bool fp32, const double *Fs,
              const double *Fb, double *Vs, double *Vb, const double *Xs,
              const double *Xb, const int *skip = nullptr);
}  // namespace gmi
