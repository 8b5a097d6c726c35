// gmi_capi.hip -- the C ABI (include/gmi.h): context, buffers, launches.
//
// Replaces rpy_mobility_solver (rpy_mobility_solver.cu:14-311).  Design:
//  * one context = one device + one stream; all device vectors resident;
//  * device layout is SoA (component-major, leading dimension padded to a
//    multiple of 64) so that every kernel load is coalesced;
//  * host copies go through per-buffer pinned staging: H2D is asynchronous
//    (the pinned region is reused only after its previous copy finished),
//    D2H completes at gmi_wait_for_device();
//  * pair launches run from precomputed work lists (target tile x source
//    split), so uneven swimmers and small problems still fill the GPU.
#include "gmi_internal.hpp"
#include "gmi_ctx.hpp"
#include "../../include/gmi.h"

#include <algorithm>
#include <tuple>
#include <cstdio>
#include <cstring>
#include <cmath>

using namespace gmi;

namespace gmi {

int fail(gmi_ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg;
  return code;
}

int hip_check(gmi_ctx *c, hipError_t e, const char *what) {
  if (e == hipSuccess) return GMI_OK;
  char buf[256];
  std::snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
  return fail(c, GMI_ERR_HIP, buf);
}

int pad64(int n) { return ((n + 63) / 64) * 64; }

// Build the work list of one block mask.

static int build_list(gmi_ctx *c, int mask, WorkList &wl, int target_blocks, int min_chunk) {
  wl.h_tiles.clear();
  std::vector<WorkItem> work;
  const bool seg_targets = mask & (kSS | kSB);
  const bool blob_targets = mask & (kBS | kBB);
  // a group: target ranges (ts0, tns segments; tb0, tnb blobs) and the
  // source ranges they see (s0, ns; b0, nb)
  struct G { int ts0, tns, tb0, tnb, s0, ns, b0, nb; };
  std::vector<G> groups;
  if (c->nsw == 1) {
    groups.push_back({0, c->Ns, 0, c->Nb, 0, c->Ns, 0, c->Nb});
  } else {
    // tiles never straddle two swimmers; coupled swimmers see every source
    for (int g = 0; g < c->nsw; ++g) {
      G q{c->seg_off[g], c->nseg[g], c->blob_off[g], c->nblob[g],
          c->seg_off[g], c->nseg[g], c->blob_off[g], c->nblob[g]};
      if (c->interact) { q.s0 = 0; q.ns = c->Ns; q.b0 = 0; q.nb = c->Nb; }
      groups.push_back(q);
    }
  }
  // total pair work, to spread splits in proportion
  long long tot_tiles = 0;
  for (const G &g : groups) {
    if (seg_targets) tot_tiles += (g.tns + kBlock - 1) / kBlock;
    if (blob_targets) tot_tiles += (g.tnb + kBlock - 1) / kBlock;
  }
  const long long target = std::max<long long>(1, target_blocks);
  const int base_split =
      int(std::max<long long>(1, (target + tot_tiles - 1) / std::max<long long>(1, tot_tiles)));
  int part0 = 0;
  for (const G &g : groups) {
    for (int sp = 0; sp < 2; ++sp) {
      const bool segt = (sp == 0);
      if (segt ? !seg_targets : !blob_targets) continue;
      const int nt_all = segt ? g.tns : g.tnb;
      const int t_off = segt ? g.ts0 : g.tb0;
      const bool use_s = segt ? (mask & kSS) : (mask & kBS);
      const bool use_b = segt ? (mask & kSB) : (mask & kBB);
      const int L = (use_s ? g.ns : 0) + (use_b ? g.nb : 0);
      const int max_split = std::max(1, (L + min_chunk - 1) / min_chunk);
      const int nsplit = std::max(1, std::min(base_split, max_split));
      for (int t = 0; t < nt_all; t += kBlock) {
        TileDesc T;
        T.species = segt ? 0 : 1;
        T.t0 = t_off + t;
        T.nt = std::min(kBlock, nt_all - t);
        T.sseg0 = g.s0; T.nsseg = g.ns;
        T.sblob0 = g.b0; T.nsblob = g.nb;
        T.nsplit = nsplit;
        T.part0 = (nsplit > 1) ? part0 : 0;
        if (nsplit > 1) part0 += nsplit;
        const int id = int(wl.h_tiles.size());
        wl.h_tiles.push_back(T);
        for (int s = 0; s < nsplit; ++s) work.push_back({id, s});
      }
    }
  }
  wl.ntiles = int(wl.h_tiles.size());
  wl.nwork = int(work.size());
  wl.nslots = part0;
  wl.nsplit_tiles = 0;
  for (const TileDesc &T : wl.h_tiles) wl.nsplit_tiles += (T.nsplit > 1);
  if (wl.d_tiles) (void)hipFree(wl.d_tiles);
  if (wl.d_work) (void)hipFree(wl.d_work);
  wl.d_tiles = nullptr; wl.d_work = nullptr;
  if (wl.ntiles == 0) return GMI_OK;
  int rc;
  if ((rc = hip_check(c, hipMalloc(&wl.d_tiles, sizeof(TileDesc) * wl.ntiles), "hipMalloc tiles")))
    return rc;
  if ((rc = hip_check(c, hipMalloc(&wl.d_work, sizeof(WorkItem) * wl.nwork), "hipMalloc work")))
    return rc;
  if ((rc = hip_check(c, hipMemcpy(wl.d_tiles, wl.h_tiles.data(), sizeof(TileDesc) * wl.ntiles,
                                   hipMemcpyHostToDevice), "upload tiles")))
    return rc;
  return hip_check(c, hipMemcpy(wl.d_work, work.data(), sizeof(WorkItem) * wl.nwork,
                                hipMemcpyHostToDevice), "upload work");
}

int rebuild_lists(gmi_ctx *c) {
  int rc;
  c->list_generation++;
  size_t slots = 0;
  for (int mask = 1; mask < 16; ++mask) {
    if ((rc = build_list(c, mask, c->lists[mask], c->target_blocks, c->min_chunk))) return rc;
    slots = std::max(slots, size_t(c->lists[mask].nslots));
  }
  if (slots > c->part_slots) {
    if (c->part) (void)hipFree(c->part);
    c->part = nullptr;
    if ((rc = hip_check(c, hipMalloc(&c->part, slots * 6 * kBlock * sizeof(double)),
                        "hipMalloc partials")))
      return rc;
    c->part_slots = slots;
  }
  return GMI_OK;
}

int prof_start(gmi_ctx *c, hipEvent_t *ev_stop) {
  *ev_stop = nullptr;
  if (!c->profiling) return GMI_OK;
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  int rc;
  if ((rc = hip_check(c, hipStreamIsCapturing(c->stream, &st), "capture status"))) return rc;
  if (st != hipStreamCaptureStatusNone) return GMI_OK;
  while (c->prof_events.size() < c->prof_used + 2) {
    hipEvent_t e;
    if ((rc = hip_check(c, hipEventCreate(&e), "event"))) return rc;
    c->prof_events.push_back(e);
  }
  if ((rc = hip_check(c, hipEventRecord(c->prof_events[c->prof_used], c->stream), "event")))
    return rc;
  *ev_stop = c->prof_events[c->prof_used + 1];
  c->prof_used += 2;
  return GMI_OK;
}

int prof_gs_start(gmi_ctx *c, int k, int n, int passes, hipEvent_t *ev_stop) {
  *ev_stop = nullptr;
  if (!c->profiling) return GMI_OK;
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  int rc;
  if ((rc = hip_check(c, hipStreamIsCapturing(c->stream, &st), "capture status"))) return rc;
  if (st != hipStreamCaptureStatusNone) return GMI_OK;
  while (c->gs_events.size() < c->gs_used + 2) {
    hipEvent_t e;
    if ((rc = hip_check(c, hipEventCreate(&e), "event"))) return rc;
    c->gs_events.push_back(e);
  }
  if ((rc = hip_check(c, hipEventRecord(c->gs_events[c->gs_used], c->stream), "event")))
    return rc;
  c->gs_recs.push_back({c->gs_used, k, n, passes});
  *ev_stop = c->gs_events[c->gs_used + 1];
  c->gs_used += 2;
  return GMI_OK;
}

int run_pairs(gmi_ctx *c, int mask, int accumulate, bool fp32,
              const double *Fs, const double *Fb, double *Vs, double *Vb,
              const double *Xs, const double *Xb, const int *skip) {
  const WorkList &wl = c->lists[mask];
  PairArgs a;
  a.Xs = Xs ? Xs : c->Xs; a.Fs = Fs ? Fs : c->Fs; a.lds = c->lds;
  a.Xb = Xb ? Xb : c->Xb; a.Fb = Fb ? Fb : c->Fb; a.ldb = c->ldb;
  a.Vs = Vs ? Vs : c->Vs; a.Vb = Vb ? Vb : c->Vb;
  a.part = c->part;
  a.tiles = wl.d_tiles;
  a.work = wl.d_work;
  a.mask = mask;
  a.accumulate = accumulate;
  a.a_s = c->as; a.a_b = c->ab; a.mu = c->mu;
  a.tt_only = c->tt_only ? 1 : 0;
  a.skip = skip;
  hipEvent_t ev_mid = nullptr;
  if (wl.nwork > 0) {
    int rc;
    if ((rc = prof_start(c, &ev_mid))) return rc;
  }
  if (c->wall)
    return hip_check(c, launch_pair_wall(a, wl.nwork, wl.nsplit_tiles ? wl.ntiles : 0, c->stream,
                                         ev_mid),
                     "wall pair launch");
  return hip_check(c, launch_pair(a, wl.nwork, wl.nsplit_tiles ? wl.ntiles : 0, fp32, c->stream,
                                  ev_mid),
                   "pair launch");
}

}  // namespace gmi

// ---- staging helpers -------------------------------------------------------

static int h2d(gmi_ctx *c, int slot, const double *host, int n, int nc,
               double *dst_soa, int ld) {
  if (!host) return fail(c, GMI_ERR_ARG, "null host pointer");
  if (n == 0) return GMI_OK;
  gmi_ctx::Stage &st = c->stage[slot];
  int rc;
  // the pinned region may still feed an earlier asynchronous copy
  if ((rc = hip_check(c, hipEventSynchronize(st.done), "stage wait"))) return rc;
  std::memcpy(st.pinned, host, sizeof(double) * size_t(n) * nc);
  if ((rc = hip_check(c, hipMemcpyAsync(st.dev, st.pinned, sizeof(double) * size_t(n) * nc,
                                        hipMemcpyHostToDevice, c->stream), "H2D")))
    return rc;
  if ((rc = hip_check(c, hipEventRecord(st.done, c->stream), "stage record"))) return rc;
  return hip_check(c, launch_aos_to_soa(dst_soa, ld, st.dev, n, nc, c->stream), "aos->soa");
}

static int flush_pending(gmi_ctx *c) {
  if (c->pending.empty()) return GMI_OK;
  int rc;
  if ((rc = hip_check(c, hipStreamSynchronize(c->stream), "stream sync"))) return rc;
  for (const gmi_ctx::Pending &p : c->pending) std::memcpy(p.dst, p.src, p.bytes);
  c->pending.clear();
  return GMI_OK;
}

static int d2h(gmi_ctx *c, int slot, double *host, int n, int nc,
               const double *src_soa, int ld) {
  if (!host) return fail(c, GMI_ERR_ARG, "null host pointer");
  if (n == 0) return GMI_OK;
  gmi_ctx::Stage &st = c->stage[slot];
  int rc;
  for (const gmi_ctx::Pending &p : c->pending)
    if (p.src == st.pinned) {
      if ((rc = flush_pending(c))) return rc;
      break;
    }
  if ((rc = hip_check(c, hipEventSynchronize(st.done), "stage wait"))) return rc;
  if ((rc = hip_check(c, launch_soa_to_aos(st.dev, src_soa, ld, n, nc, c->stream), "soa->aos")))
    return rc;
  const size_t bytes = sizeof(double) * size_t(n) * nc;
  if ((rc = hip_check(c, hipMemcpyAsync(st.pinned, st.dev, bytes, hipMemcpyDeviceToHost,
                                        c->stream), "D2H")))
    return rc;
  if ((rc = hip_check(c, hipEventRecord(st.done, c->stream), "stage record"))) return rc;
  c->pending.push_back({host, st.pinned, bytes});
  return GMI_OK;
}

// ---- C ABI -------------------------------------------------------------------

extern "C" {

int gmi_device_count(int *count) {
  if (!count) return GMI_ERR_ARG;
  *count = 0;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return GMI_OK;  // no runtime / no device: zero devices
  }
  *count = n;
  return GMI_OK;
}

int gmi_create(gmi_ctx **out, const gmi_params *p) {
  if (!out || !p) return GMI_ERR_ARG;
  *out = nullptr;
  if (p->n_swimmers < 1 || !p->nseg_per_swimmer || !p->nblob_per_swimmer ||
      p->nseg_per_filament < 1 || !(p->mu > 0) || !(p->a_seg > 0) || !(p->a_blob > 0) ||
      (p->precision != GMI_FP64 && p->precision != GMI_FP32_KAHAN))
    return GMI_ERR_ARG;
  gmi_ctx *c = new gmi_ctx();
  c->device = p->device;
  c->nsw = p->n_swimmers;
  c->nseg.assign(p->nseg_per_swimmer, p->nseg_per_swimmer + p->n_swimmers);
  c->nblob.assign(p->nblob_per_swimmer, p->nblob_per_swimmer + p->n_swimmers);
  c->Ns = 0; c->Nb = 0;
  for (int g = 0; g < c->nsw; ++g) {
    if (c->nseg[g] < 0 || c->nblob[g] < 0 || c->nseg[g] % p->nseg_per_filament) {
      delete c;
      return GMI_ERR_ARG;
    }
    c->seg_off.push_back(c->Ns);
    c->blob_off.push_back(c->Nb);
    c->Ns += c->nseg[g];
    c->Nb += c->nblob[g];
  }
  c->nseg_fil = p->nseg_per_filament;
  c->interact = p->swimmers_interact != 0;
  c->mu = p->mu; c->as = p->a_seg; c->ab = p->a_blob;
  c->precision = p->precision;
  c->lds = pad64(std::max(c->Ns, 1));
  c->ldb = pad64(std::max(c->Nb, 1));
  *out = c;
  int rc;
  if ((rc = hip_check(c, hipSetDevice(c->device), "hipSetDevice"))) return rc;
  if ((rc = hip_check(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking),
                      "stream")))
    return rc;
  if ((rc = hip_check(c, hipEventCreate(&c->ev0), "event"))) return rc;
  if ((rc = hip_check(c, hipEventCreate(&c->ev1), "event"))) return rc;
  auto dalloc = [&](double **ptr, size_t n) {
    return hip_check(c, hipMalloc(ptr, sizeof(double) * std::max<size_t>(n, 1)), "hipMalloc");
  };
  if ((rc = dalloc(&c->Xs, 3 * size_t(c->lds)))) return rc;
  if ((rc = dalloc(&c->Fs, 6 * size_t(c->lds)))) return rc;
  if ((rc = dalloc(&c->Vs, 6 * size_t(c->lds)))) return rc;
  if ((rc = dalloc(&c->Xb, 3 * size_t(c->ldb)))) return rc;
  if ((rc = dalloc(&c->Fb, 3 * size_t(c->ldb)))) return rc;
  if ((rc = dalloc(&c->Vb, 3 * size_t(c->ldb)))) return rc;
  if ((rc = dalloc(&c->Frep, 3 * size_t(c->ldb)))) return rc;
  {
    // the barrier's cell grid: one to two buckets per particle, cells as
    // wide as the largest contact distance 1.1 (a_i + a_j)
    CellGrid &G = c->cells;
    const int N = c->Ns + c->Nb;
    G.H = kScanChunk;
    while (G.H < N) G.H *= 2;
    G.inv_h = 1.0 / (1.1 * 2.0 * std::max(c->as, c->ab));
    std::vector<int> grp(std::max(N, 1), 0);
    if (!c->interact)
      for (int g = 0; g < c->nsw; ++g) {
        for (int k = 0; k < c->nseg[g]; ++k) grp[c->seg_off[g] + k] = g;
        for (int k = 0; k < c->nblob[g]; ++k) grp[c->Ns + c->blob_off[g] + k] = g;
      }
    int *gp = nullptr;
    auto ialloc = [&](int **ptr, size_t n) {
      return hip_check(c, hipMalloc(ptr, sizeof(int) * std::max<size_t>(n, 1)), "hipMalloc");
    };
    if ((rc = ialloc(&G.cnt, G.H)) || (rc = ialloc(&G.fill, G.H)) ||
        (rc = ialloc(&G.start, G.H)) || (rc = ialloc(&G.boff, G.H / kScanChunk)) ||
        (rc = ialloc(&G.ent, N)) || (rc = ialloc(&G.arr, N)) || (rc = ialloc(&G.bucket, N)) ||
        (rc = ialloc(&gp, N)) ||
        (rc = hip_check(c, hipMalloc(&G.key, sizeof(int4) * std::max(N, 1)), "hipMalloc")) ||
        (rc = hip_check(c, hipMalloc(&G.rk, sizeof(int4) * std::max(N, 1)), "hipMalloc")) ||
        (rc = dalloc(&G.rx, N)) || (rc = dalloc(&G.ry, N)) || (rc = dalloc(&G.rz, N)))
      return rc;
    G.grp = gp;
    if ((rc = hip_check(c, hipMemset(G.cnt, 0, sizeof(int) * G.H), "memset")) ||
        (rc = hip_check(c, hipMemset(G.fill, 0, sizeof(int) * G.H), "memset")) ||
        (rc = hip_check(c, hipMemcpy(gp, grp.data(), sizeof(int) * grp.size(),
                                     hipMemcpyHostToDevice), "upload groups")))
      return rc;
  }
  // zero every vector once so that partial uploads never expose garbage
  double *bufs[] = {c->Xs, c->Fs, c->Vs, c->Xb, c->Fb, c->Vb, c->Frep};
  size_t lens[] = {3u * c->lds, 6u * c->lds, 6u * c->lds, 3u * c->ldb,
                   3u * c->ldb, 3u * c->ldb, 3u * c->ldb};
  for (int k = 0; k < 7; ++k)
    if ((rc = hip_check(c, hipMemset(bufs[k], 0, sizeof(double) * lens[k]), "memset")))
      return rc;
  const size_t stage_n[gmi_ctx::kNumStage] = {
      3u * c->Ns, 6u * c->Ns, 3u * c->Nb, 3u * c->Nb, 3u * c->Nb, 3u * c->Nb, 6u * c->Ns};
  for (int k = 0; k < gmi_ctx::kNumStage; ++k) {
    gmi_ctx::Stage &st = c->stage[k];
    const size_t bytes = sizeof(double) * std::max<size_t>(stage_n[k], 1);
    if ((rc = hip_check(c, hipHostMalloc((void **)&st.pinned, bytes, hipHostMallocDefault),
                        "hipHostMalloc")))
      return rc;
    if ((rc = hip_check(c, hipMalloc(&st.dev, bytes), "hipMalloc stage"))) return rc;
    if ((rc = hip_check(c, hipEventCreateWithFlags(&st.done, hipEventDisableTiming), "event")))
      return rc;
    if ((rc = hip_check(c, hipEventRecord(st.done, c->stream), "event record"))) return rc;
  }
  if ((rc = rebuild_lists(c))) return rc;
  return GMI_OK;
}

void gmi_destroy(gmi_ctx *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  c->release_extra();
  double *bufs[] = {c->Xs, c->Fs, c->Vs, c->Xb, c->Fb, c->Vb, c->Frep, c->part};
  for (double *b : bufs)
    if (b) (void)hipFree(b);
  for (auto &kv : c->lists) {
    if (kv.second.d_tiles) (void)hipFree(kv.second.d_tiles);
    if (kv.second.d_work) (void)hipFree(kv.second.d_work);
  }
  {
    CellGrid &G = c->cells;
    void *cb[] = {G.cnt, G.fill, G.start, G.boff, G.ent, G.arr, G.bucket, G.key, G.rk, G.rx, G.ry,
                  G.rz, const_cast<int *>(G.grp)};
    for (void *b : cb)
      if (b) (void)hipFree(b);
  }
  for (auto &st : c->stage) {
    if (st.pinned) (void)hipHostFree(st.pinned);
    if (st.dev) (void)hipFree(st.dev);
    if (st.done) (void)hipEventDestroy(st.done);
  }
  for (hipEvent_t e : c->prof_events) (void)hipEventDestroy(e);
  for (hipEvent_t e : c->gs_events) (void)hipEventDestroy(e);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char *gmi_last_error(const gmi_ctx *c) { return c ? c->err.c_str() : "null context"; }

int gmi_set_precision(gmi_ctx *c, int precision) {
  if (!c) return GMI_ERR_ARG;
  if (precision != GMI_FP64 && precision != GMI_FP32_KAHAN)
    return fail(c, GMI_ERR_ARG, "unknown precision");
  c->precision = precision;
  return GMI_OK;
}

int gmi_num_segments(const gmi_ctx *c) { return c ? c->Ns : GMI_ERR_ARG; }
int gmi_num_blobs(const gmi_ctx *c) { return c ? c->Nb : GMI_ERR_ARG; }

int gmi_copy_segment_positions_to_device(gmi_ctx *c, const double *x) {
  if (!c) return GMI_ERR_ARG;
  return h2d(c, gmi_ctx::kStXs, x, c->Ns, 3, c->Xs, c->lds);
}
int gmi_copy_segment_forces_to_device(gmi_ctx *c, const double *f) {
  if (!c) return GMI_ERR_ARG;
  return h2d(c, gmi_ctx::kStFs, f, c->Ns, 6, c->Fs, c->lds);
}
int gmi_copy_blob_positions_to_device(gmi_ctx *c, const double *x) {
  if (!c) return GMI_ERR_ARG;
  return h2d(c, gmi_ctx::kStXb, x, c->Nb, 3, c->Xb, c->ldb);
}
int gmi_copy_blob_forces_to_device(gmi_ctx *c, const double *f) {
  if (!c) return GMI_ERR_ARG;
  return h2d(c, gmi_ctx::kStFb, f, c->Nb, 3, c->Fb, c->ldb);
}
int gmi_copy_interparticle_blob_forces_to_host(gmi_ctx *c, double *f) {
  if (!c) return GMI_ERR_ARG;
  return d2h(c, gmi_ctx::kStRep, f, c->Nb, 3, c->Frep, c->ldb);
}
int gmi_copy_blob_velocities_to_host(gmi_ctx *c, double *v) {
  if (!c) return GMI_ERR_ARG;
  return d2h(c, gmi_ctx::kStVb, v, c->Nb, 3, c->Vb, c->ldb);
}
int gmi_copy_segment_velocities_to_host(gmi_ctx *c, double *v) {
  if (!c) return GMI_ERR_ARG;
  return d2h(c, gmi_ctx::kStVs, v, c->Ns, 6, c->Vs, c->lds);
}

int gmi_apply_interparticle_forces(gmi_ctx *c, double K) {
  if (!c) return GMI_ERR_ARG;
  if (!std::isfinite(K)) return fail(c, GMI_ERR_ARG, "barrier K not finite");
  return hip_check(c, launch_barrier(c->Fs, c->lds, c->Frep, c->ldb, c->Xs, c->Xb, c->Ns, c->Nb,
                                     c->cells, c->nseg_fil, c->as, c->ab, K, c->stream),
                   "barrier launch");
}

int gmi_wait_for_device(gmi_ctx *c) {
  if (!c) return GMI_ERR_ARG;
  int rc;
  if ((rc = hip_check(c, hipStreamSynchronize(c->stream), "stream sync"))) return rc;
  return flush_pending(c);
}

int gmi_set_wall(gmi_ctx *c, int wall) {
  if (!c) return GMI_ERR_ARG;
  if (wall && c->Nb > 0)
    return fail(c, GMI_ERR_STATE, "the wall-bounded mobility has segments only (no blobs)");
  if (c->extra && (wall != 0) != c->wall)
    return fail(c, GMI_ERR_STATE, "the wall setting is fixed while a time stepper exists");
  c->wall = wall != 0;
  return GMI_OK;
}

int gmi_evaluate_segment_segment_mobility(gmi_ctx *c) {
  if (!c) return GMI_ERR_ARG;
  return run_pairs(c, kSS, 0, c->precision == GMI_FP32_KAHAN, nullptr, nullptr, nullptr,
                   nullptr, nullptr, nullptr);
}
int gmi_evaluate_segment_blob_mobility(gmi_ctx *c) {
  if (!c) return GMI_ERR_ARG;
  return run_pairs(c, kSB, 1, c->precision == GMI_FP32_KAHAN, nullptr, nullptr, nullptr,
                   nullptr, nullptr, nullptr);
}
int gmi_evaluate_blob_blob_mobility(gmi_ctx *c) {
  if (!c) return GMI_ERR_ARG;
  return run_pairs(c, kBB, 0, c->precision == GMI_FP32_KAHAN, nullptr, nullptr, nullptr,
                   nullptr, nullptr, nullptr);
}
int gmi_evaluate_blob_segment_mobility(gmi_ctx *c) {
  if (!c) return GMI_ERR_ARG;
  return run_pairs(c, kBS, 0, c->precision == GMI_FP32_KAHAN, nullptr, nullptr, nullptr,
                   nullptr, nullptr, nullptr);
}
int gmi_evaluate_grand_mobility(gmi_ctx *c) {
  if (!c) return GMI_ERR_ARG;
  return run_pairs(c, kSS | kSB | kBS | kBB, 0, c->precision == GMI_FP32_KAHAN, nullptr,
                   nullptr, nullptr, nullptr, nullptr, nullptr);
}

int gmi_timer_start(gmi_ctx *c) {
  if (!c) return GMI_ERR_ARG;
  return hip_check(c, hipEventRecord(c->ev0, c->stream), "event record");
}

int gmi_timer_stop(gmi_ctx *c, double *ms) {
  if (!c || !ms) return GMI_ERR_ARG;
  int rc;
  if ((rc = hip_check(c, hipEventRecord(c->ev1, c->stream), "event record"))) return rc;
  if ((rc = hip_check(c, hipEventSynchronize(c->ev1), "event sync"))) return rc;
  float t = 0.f;
  if ((rc = hip_check(c, hipEventElapsedTime(&t, c->ev0, c->ev1), "elapsed"))) return rc;
  *ms = t;
  return GMI_OK;
}

int gmi_profile_begin(gmi_ctx *c) {
  if (!c) return GMI_ERR_ARG;
  c->profiling = true;
  c->prof_used = 0;
  c->gs_used = 0;
  c->gs_recs.clear();
  return GMI_OK;
}

int gmi_profile_krylov(gmi_ctx *c, double *gs_ms, double *gs_bytes, int *steps) {
  if (!c || !gs_ms || !gs_bytes || !steps) return GMI_ERR_ARG;
  int rc;
  if ((rc = hip_check(c, hipStreamSynchronize(c->stream), "stream sync"))) return rc;
  double tot = 0.0, bytes = 0.0;
  for (const gmi_ctx::GsRec &g : c->gs_recs) {
    float t = 0.f;
    if ((rc = hip_check(c, hipEventElapsedTime(&t, c->gs_events[g.ev], c->gs_events[g.ev + 1]),
                        "elapsed")))
      return rc;
    tot += t;
    // per Gram-Schmidt pass: k columns and w read, w written
    bytes += double(g.passes) * 8.0 * double(g.n) * double(g.k + 2);
  }
  *gs_ms = tot;
  *gs_bytes = bytes;
  *steps = int(c->gs_recs.size());
  c->gs_recs.clear();
  c->gs_used = 0;
  return GMI_OK;
}

int gmi_profile_end(gmi_ctx *c, double *pair_ms, int *launches) {
  if (!c || !pair_ms || !launches) return GMI_ERR_ARG;
  int rc;
  c->profiling = false;
  if ((rc = hip_check(c, hipStreamSynchronize(c->stream), "stream sync"))) return rc;
  double tot = 0.0;
  for (size_t k = 0; k + 1 < c->prof_used; k += 2) {
    float t = 0.f;
    if ((rc = hip_check(c, hipEventElapsedTime(&t, c->prof_events[k], c->prof_events[k + 1]),
                        "elapsed")))
      return rc;
    tot += t;
  }
  *pair_ms = tot;
  *launches = int(c->prof_used / 2);
  c->prof_used = 0;
  return GMI_OK;
}

int gmi_set_launch_shape(gmi_ctx *c, int target_blocks, int min_chunk);

// Launch-shape autotuner: times the grand apply (all four blocks) on the
// context's current device data for each candidate (target workgroups,
// minimum source chunk) and keeps the fastest.  Results are cached per
// problem shape for the life of the process.
namespace {
struct TuneKey {
  int Ns, Nb, nsw, interact, precision, mask;
  bool operator<(const TuneKey &o) const {
    return std::tie(Ns, Nb, nsw, interact, precision, mask) <
           std::tie(o.Ns, o.Nb, o.nsw, o.interact, o.precision, o.mask);
  }
};
std::map<TuneKey, std::pair<int, int>> g_tuned;

// Times `reps` launches of the `mask` blocks for each launch shape of a grid
// and keeps the fastest (cached per sizes, precision and mask).
int tune_shape(gmi_ctx *c, int mask, int reps, int *best_target, int *best_chunk) {
  const TuneKey key{c->Ns, c->Nb, c->nsw, int(c->interact), c->precision, mask};
  int rc;
  auto it = g_tuned.find(key);
  if (it == g_tuned.end()) {
    const int targets[] = {256, 512, 1024, 2048, 4096, 8192};
    const int chunks[] = {64, 128, 256, 512, 1024, 2048};
    double best = 1e300;
    std::pair<int, int> arg{c->target_blocks, c->min_chunk};
    hipEvent_t e0, e1;
    if ((rc = hip_check(c, hipEventCreate(&e0), "event"))) return rc;
    if ((rc = hip_check(c, hipEventCreate(&e1), "event"))) return rc;
    const bool fp32 = c->precision == GMI_FP32_KAHAN;
    for (int tb : targets)
      for (int ch : chunks) {
        if ((rc = gmi_set_launch_shape(c, tb, ch))) return rc;
        if ((rc = run_pairs(c, mask, 0, fp32, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr)))
          return rc;
        (void)hipEventRecord(e0, c->stream);
        for (int r = 0; r < reps; r++)
          if ((rc = run_pairs(c, mask, 0, fp32, nullptr, nullptr, nullptr, nullptr, nullptr,
                              nullptr)))
            return rc;
        (void)hipEventRecord(e1, c->stream);
        if ((rc = hip_check(c, hipEventSynchronize(e1), "tune sync"))) return rc;
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (ms < best) { best = ms; arg = {tb, ch}; }
      }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    it = g_tuned.emplace(key, arg).first;
  }
  if ((rc = gmi_set_launch_shape(c, it->second.first, it->second.second))) return rc;
  if (best_target) *best_target = it->second.first;
  if (best_chunk) *best_chunk = it->second.second;
  return GMI_OK;
}
}  // namespace

int gmi_autotune(gmi_ctx *c, int reps, int *best_target, int *best_chunk) {
  if (!c || reps < 1) return GMI_ERR_ARG;
  return tune_shape(c, kSS | kSB | kBS | kBB, reps, best_target, best_chunk);
}

int gmi_autotune_blocks(gmi_ctx *c, int mask, int reps, int *best_target, int *best_chunk) {
  if (!c) return GMI_ERR_ARG;
  if (reps < 1 || mask < 1 || mask > 15) return fail(c, GMI_ERR_ARG, "mask in [1, 15], reps >= 1");
  return tune_shape(c, mask, reps, best_target, best_chunk);
}

int gmi_set_launch_shape(gmi_ctx *c, int target_blocks, int min_chunk) {
  if (!c) return GMI_ERR_ARG;
  if (target_blocks < 1 || min_chunk < 1) return fail(c, GMI_ERR_ARG, "bad launch shape");
  int rc;
  if ((rc = hip_check(c, hipStreamSynchronize(c->stream), "stream sync"))) return rc;
  c->target_blocks = target_blocks;
  c->min_chunk = min_chunk;
  return rebuild_lists(c);
}

}  // extern "C"
