// gmi_io.hip -- simulation output and checkpoint files in the reference's
// text formats (include/gmi_sim.h).  Host code only: the device state is
// copied to the host, formatted as the reference's streams format it
// (std::scientific, precision 10 -> "%.10e"; default stream format -> "%g"),
// and written; backups are parsed and uploaded back.
//
// Files (SIMULATION_NAME = `name`, config.hpp:208-215):
//   <name>.par                   cilia_sim_main.cpp:79-116
//   <name>_seg_states.dat        cilia_sim_main.cpp:346-363, filament.cpp:1577-1591
//   <name>_body_states.dat       cilia_sim_main.cpp:346-363, rigid_body.cpp:408-412
//   <name>_body_vels.dat, _blob_forces.dat, _seg_vels.dat, _seg_forces.dat
//                                mobility_solver.cpp:1775-1895 (GMRES build)
//   <name>.backup                cilia_sim_main.cpp:368-381, swimmer.cpp:914-925,
//                                filament.cpp:1593-1650, rigid_body.cpp:414-432
//   <name>_blob_references.dat, <name>_fil_references.dat
//                                swimmer.cpp:885-900, rigid_body.cpp:383-400
// plus <name>.backup_gmi (no reference counterpart): the same state at full
// precision with the GMRES warm start and the learned warm start's lag
// features, so that a resumed run repeats the original bit for bit.
#include "gmi_ctx.hpp"
#include "gmi_sim_state.hpp"

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

namespace gmi {

// learned warm start lag buffer access (gmi_learn.hip)
int learn_lags(gmi_ctx *c, SimHost *s, std::vector<float> *get, const std::vector<float> *set);

namespace {

SimHost *sim_of(gmi_ctx *c) {
  return (c && c->extra_kind == kExtraSim) ? static_cast<SimHost *>(c->extra) : nullptr;
}

// Host copy of the device state, SoA as on the device.
struct HostState {
  std::vector<double> X, Xm1, Xm2, Q, U, Um1, lam, teth, clamp, body, Vs, Fs, x;
};

template <typename T>
int d2h_vec(gmi_ctx *c, std::vector<T> &v, const T *src, size_t n) {
  v.resize(n);
  if (n == 0) return GMI_OK;
  // null-stream copies do not wait for the context's (non-blocking) stream
  int rc;
  if ((rc = hip_check(c, hipStreamSynchronize(c->stream), "sync"))) return rc;
  return hip_check(c, hipMemcpy(v.data(), src, sizeof(T) * n, hipMemcpyDeviceToHost), "state D2H");
}
template <typename T>
int h2d_vec(gmi_ctx *c, T *dst, const std::vector<T> &v) {
  if (v.empty()) return GMI_OK;
  int rc;
  if ((rc = hip_check(c, hipStreamSynchronize(c->stream), "sync"))) return rc;
  return hip_check(c, hipMemcpy(dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice),
                   "state H2D");
}

int download(gmi_ctx *c, SimHost *s, HostState &h) {
  const SimDev &d = s->d;
  const size_t ls = size_t(d.lds);
  int rc;
  if ((rc = hip_check(c, hipStreamSynchronize(c->stream), "sync"))) return rc;
  if ((rc = d2h_vec(c, h.X, static_cast<const double *>(d.X), 3 * ls))) return rc;
  if ((rc = d2h_vec(c, h.Xm1, static_cast<const double *>(d.Xm1), 3 * ls))) return rc;
  if ((rc = d2h_vec(c, h.Xm2, static_cast<const double *>(d.Xm2), 3 * ls))) return rc;
  if ((rc = d2h_vec(c, h.Q, static_cast<const double *>(d.Q), 4 * ls))) return rc;
  if ((rc = d2h_vec(c, h.U, static_cast<const double *>(d.U), 3 * ls))) return rc;
  if ((rc = d2h_vec(c, h.Um1, static_cast<const double *>(d.Um1), 3 * ls))) return rc;
  if ((rc = d2h_vec(c, h.lam, static_cast<const double *>(d.lam), size_t(d.Nf) * 9 * d.S))) return rc;
  if ((rc = d2h_vec(c, h.teth, static_cast<const double *>(d.teth), 9 * size_t(d.Nf)))) return rc;
  if ((rc = d2h_vec(c, h.clamp, static_cast<const double *>(d.clamp), 9 * size_t(d.Nf)))) return rc;
  if ((rc = d2h_vec(c, h.body, static_cast<const double *>(d.body), size_t(d.Nsw) * BREC))) return rc;
  if ((rc = d2h_vec(c, h.Vs, static_cast<const double *>(c->Vs), 6 * ls))) return rc;
  if ((rc = d2h_vec(c, h.Fs, static_cast<const double *>(c->Fs), 6 * ls))) return rc;
  return d2h_vec(c, h.x, static_cast<const double *>(s->x), size_t(d.npad));
}

int upload(gmi_ctx *c, SimHost *s, const HostState &h) {
  const SimDev &d = s->d;
  int rc;
  // the copies below run on the null stream, which does not order with the
  // context's non-blocking stream
  if ((rc = hip_check(c, hipStreamSynchronize(c->stream), "sync"))) return rc;
  if ((rc = h2d_vec(c, d.X, h.X))) return rc;
  if ((rc = h2d_vec(c, d.Xm1, h.Xm1))) return rc;
  if ((rc = h2d_vec(c, d.Xm2, h.Xm2))) return rc;
  if ((rc = h2d_vec(c, d.Q, h.Q))) return rc;
  if ((rc = h2d_vec(c, d.Qm1, h.Q))) return rc;  // q_m1 = q (filament.cpp:116)
  if ((rc = h2d_vec(c, d.U, h.U))) return rc;
  if ((rc = h2d_vec(c, d.Um1, h.Um1))) return rc;
  if ((rc = h2d_vec(c, d.lam, h.lam))) return rc;
  if ((rc = h2d_vec(c, d.teth, h.teth))) return rc;
  if ((rc = h2d_vec(c, d.clamp, h.clamp))) return rc;
  if ((rc = h2d_vec(c, d.body, h.body))) return rc;
  if ((rc = h2d_vec(c, c->Vs, h.Vs))) return rc;
  return h2d_vec(c, s->x, h.x);
}

// Text accumulated in memory, then written in one call.
struct Text {
  std::string s;
  void sci(double v) {  // std::scientific << std::setprecision(10) << v << " "
    char b[40];
    std::snprintf(b, sizeof b, "%.10e ", v);
    s += b;
  }
  void full(double v) {  // round-trip precision
    char b[40];
    std::snprintf(b, sizeof b, "%.17e ", v);
    s += b;
  }
  void gen(double v) {  // default stream format (precision 6)
    char b[40];
    std::snprintf(b, sizeof b, "%g", v);
    s += b;
  }
  void integer(long v) { s += std::to_string(v); }
  void raw(const char *t) { s += t; }
};

int write_file(gmi_ctx *c, const std::string &path, const std::string &text, bool append) {
  FILE *f = std::fopen(path.c_str(), append ? "ab" : "wb");
  if (!f) return fail(c, GMI_ERR_ARG, ("cannot open " + path).c_str());
  const size_t n = std::fwrite(text.data(), 1, text.size(), f);
  const int e = std::fclose(f);
  if (n != text.size() || e != 0) return fail(c, GMI_ERR_STATE, ("write failed: " + path).c_str());
  return GMI_OK;
}

int read_file(const std::string &path, std::string *out) {
  FILE *f = std::fopen(path.c_str(), "rb");
  if (!f) return -1;
  char buf[1 << 16];
  size_t n;
  out->clear();
  while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) out->append(buf, n);
  std::fclose(f);
  return 0;
}

// whitespace-separated numbers
struct Reader {
  const char *p, *end;
  bool ok = true;
  explicit Reader(const std::string &s) : p(s.data()), end(s.data() + s.size()) {}
  double num() {
    while (p < end && (*p == ' ' || *p == '\n' || *p == '\r' || *p == '\t')) p++;
    if (p >= end) { ok = false; return 0.0; }
    char *q = nullptr;
    errno = 0;
    const double v = std::strtod(p, &q);
    if (q == p || errno == ERANGE) { ok = false; return 0.0; }
    p = q;
    return v;
  }
  bool at_end() {
    while (p < end && (*p == ' ' || *p == '\n' || *p == '\r' || *p == '\t')) p++;
    return p >= end;
  }
};

// index helpers into the host state
inline size_t seg_g(const SimHost *s, int f, int k) { return size_t(f) * s->d.S + k; }

// one filament's backup record (filament.cpp:1629-1648), appended or parsed
template <typename F>
void filament_record(SimHost *s, HostState &h, int f, F &&io) {
  const int S = s->d.S;
  const size_t ls = size_t(s->d.lds);
  for (int k = 0; k < S; k++) {
    const size_t g = seg_g(s, f, k);
    for (int c = 0; c < 4; c++) io(h.Q[c * ls + g]);
    for (int c = 0; c < 3; c++) io(h.X[c * ls + g]);
    for (int c = 0; c < 3; c++) io(h.Xm1[c * ls + g]);
    for (int c = 0; c < 3; c++) io(h.Xm2[c * ls + g]);
    for (int c = 0; c < 3; c++) io(h.U[c * ls + g]);
    for (int c = 0; c < 3; c++) io(h.Um1[c * ls + g]);
  }
  for (int q = 0; q < 9; q++) io(h.teth[9 * size_t(f) + q]);   // value, m1, m2
  for (int q = 0; q < 9; q++) io(h.clamp[9 * size_t(f) + q]);
  double *L = h.lam.data() + size_t(f) * 9 * S;
  const int n3 = 3 * (S - 1);
  for (int n = 0; n < S - 1; n++)
    for (int hist = 0; hist < 3; hist++)
      for (int c = 0; c < 3; c++) io(L[hist * n3 + 3 * n + c]);
}

// one body's backup record (rigid_body.cpp:414-424)
template <typename F>
void body_record(HostState &h, int n, F &&io) {
  double *b = h.body.data() + size_t(n) * BREC;
  for (int c = 0; c < 4; c++) io(b[BQ + c]);
  for (int c = 0; c < 3; c++) io(b[BX + c]);
  for (int c = 0; c < 3; c++) io(b[BXM1 + c]);
  for (int c = 0; c < 3; c++) io(b[BXM2 + c]);
  for (int c = 0; c < 3; c++) io(b[BU + c]);
  for (int c = 0; c < 3; c++) io(b[BUM1 + c]);
}

// whole backup in the reference's order: per swimmer, its filaments, then
// its body (swimmer.cpp:914-925)
template <typename F>
void backup_records(SimHost *s, HostState &h, F &&io) {
  const SimDev &d = s->d;
  int f = 0;
  for (int n = 0; n < d.Nsw; n++) {
    for (int i = 0; i < s->nfil[n]; i++, f++) filament_record(s, h, f, io);
    body_record(h, n, io);
  }
}

constexpr const char *kExtMagic = "gmi-backup-1";

}  // namespace
}  // namespace gmi

using namespace gmi;

extern "C" {

int gmi_sim_write_par(gmi_ctx *c, const char *name, int total_time_steps, int plot_frequency) {
  SimHost *s = sim_of(c);
  if (!s) return c ? fail(c, GMI_ERR_STATE, "no sim") : GMI_ERR_ARG;
  if (!name) return fail(c, GMI_ERR_ARG, "no name");
  for (int n = 1; n < s->d.Nsw; n++)
    if (s->nfil[n] != s->nfil[0] || c->nblob[n] != c->nblob[0])
      return fail(c, GMI_ERR_ARG, "the .par format describes identical swimmers");
  Text t;
  auto line = [&](double v, const char *tag) {
    t.gen(v);
    t.raw(" %% ");
    t.raw(tag);
    t.raw("\n");
  };
  line(s->nfil[0], "NFIL");
  line(s->d.S, "NSEG");
  line(s->d.Nsw, "NSWIM");
  line(c->nblob[0], "NBLOB");
  line(c->mu, "MU");
  line(c->as, "RSEG");
  line(c->ab, "RBLOB");
  line(s->p.kb, "KB");
  line(s->p.kt, "KT");
  line(s->p.tol, "TOL");
  line(total_time_steps, "TOTAL_TIME_STEPS");
  line(s->p.dt, "DT");
  line(plot_frequency, "PLOT_FREQUENCY_IN_STEPS");
  return write_file(c, std::string(name) + ".par", t.s, false);
}

int gmi_sim_write_reference_positions(gmi_ctx *c, const char *name) {
  SimHost *s = sim_of(c);
  if (!s) return c ? fail(c, GMI_ERR_STATE, "no sim") : GMI_ERR_ARG;
  if (!name) return fail(c, GMI_ERR_ARG, "no name");
  // swimmers[0] only (cilia_sim_main.cpp:155)
  Text b, f;
  for (int k = 0; k < 3 * c->nblob[0]; k++) {
    b.gen(s->p.blob_refs[k]);
    b.raw(" ");
  }
  b.raw("\n");
  for (int k = 0; k < 3 * s->nfil[0]; k++) {
    f.gen(s->p.fil_refs[k]);
    f.raw(" ");
  }
  f.raw("\n");
  int rc;
  if ((rc = write_file(c, std::string(name) + "_blob_references.dat", b.s, false))) return rc;
  return write_file(c, std::string(name) + "_fil_references.dat", f.s, false);
}

int gmi_sim_write_data(gmi_ctx *c, int nt, const char *name) {
  SimHost *s = sim_of(c);
  if (!s) return c ? fail(c, GMI_ERR_STATE, "no sim") : GMI_ERR_ARG;
  if (!name) return fail(c, GMI_ERR_ARG, "no name");
  HostState h;
  int rc;
  if ((rc = download(c, s, h))) return rc;
  const SimDev &d = s->d;
  const size_t ls = size_t(d.lds), lb = size_t(d.ldb);
  Text seg, body, bvel, bfor, svel, sfor;
  seg.integer(nt + 1); seg.raw(" ");
  body.integer(nt + 1); body.raw(" ");
  for (Text *t : {&bvel, &bfor, &svel, &sfor}) { t->integer(nt); t->raw(" "); }
  int f = 0;
  for (int n = 0; n < d.Nsw; n++) {
    const double *b = h.body.data() + size_t(n) * BREC;
    for (int k = 0; k < 3; k++) body.sci(b[BX + k]);
    for (int k = 0; k < 4; k++) body.sci(b[BQ + k]);
    for (int i = 0; i < s->nfil[n]; i++, f++)
      for (int k = 0; k < d.S; k++)
        for (int q = 0; q < 4; q++) seg.sci(h.Q[q * ls + seg_g(s, f, k)]);
    for (int k = 0; k < 6; k++) bvel.sci(h.x[3 * lb + 6 * n + k]);
    for (int m = c->blob_off[n]; m < c->blob_off[n] + c->nblob[n]; m++)
      for (int k = 0; k < 3; k++) bfor.sci(h.x[k * lb + m]);
    for (int g = c->seg_off[n]; g < c->seg_off[n] + c->nseg[n]; g++)
      for (int k = 0; k < 6; k++) {
        svel.sci(h.Vs[k * ls + g]);
      }
    for (int g = c->seg_off[n]; g < c->seg_off[n] + c->nseg[n]; g++)
      for (int k = 0; k < 6; k++) sfor.sci(h.Fs[k * ls + g]);
  }
  for (Text *t : {&seg, &body, &bvel, &bfor, &svel, &sfor}) t->raw("\n");
  const std::string nm(name);
  if ((rc = write_file(c, nm + "_seg_states.dat", seg.s, true))) return rc;
  if ((rc = write_file(c, nm + "_body_states.dat", body.s, true))) return rc;
  if ((rc = write_file(c, nm + "_body_vels.dat", bvel.s, true))) return rc;
  if ((rc = write_file(c, nm + "_blob_forces.dat", bfor.s, true))) return rc;
  if ((rc = write_file(c, nm + "_seg_vels.dat", svel.s, true))) return rc;
  return write_file(c, nm + "_seg_forces.dat", sfor.s, true);
}

int gmi_sim_write_backup(gmi_ctx *c, int nt, const char *name, int extended) {
  SimHost *s = sim_of(c);
  if (!s) return c ? fail(c, GMI_ERR_STATE, "no sim") : GMI_ERR_ARG;
  if (!name) return fail(c, GMI_ERR_ARG, "no name");
  HostState h;
  int rc;
  if ((rc = download(c, s, h))) return rc;
  Text t;
  t.integer(nt + 1);
  t.raw(" ");
  backup_records(s, h, [&](double &v) { t.sci(v); });
  t.raw("\n");
  const std::string nm(name);
  if ((rc = write_file(c, nm + ".backup", t.s, false))) return rc;
  if (!extended) return GMI_OK;
  // full precision, plus what the reference format leaves out
  Text e;
  std::vector<float> lags;
  if ((rc = learn_lags(c, s, &lags, nullptr))) return rc;
  e.raw(kExtMagic);
  e.raw(" ");
  for (long v : {long(nt + 1), long(s->d.Ns), long(s->d.Nf), long(s->d.Nsw), long(s->d.npad),
                 long(lags.size())}) {
    e.integer(v);
    e.raw(" ");
  }
  e.raw("\n");
  backup_records(s, h, [&](double &v) { e.full(v); });
  e.raw("\n");
  const SimDev &d = s->d;
  const size_t ls = size_t(d.lds);
  for (int k = 0; k < 6; k++)
    for (int g = 0; g < d.Ns; g++) e.full(h.Vs[k * ls + g]);
  e.raw("\n");
  for (double v : h.x) e.full(v);
  e.raw("\n");
  e.full(s->iters_avg);
  e.raw("\n");
  for (float v : lags) {
    char b[32];
    std::snprintf(b, sizeof b, "%.9e ", double(v));
    e.raw(b);
  }
  e.raw("\n");
  return write_file(c, nm + ".backup_gmi", e.s, false);
}

int gmi_sim_read_backup(gmi_ctx *c, const char *name, int *nt_of_backup, int *exact) {
  SimHost *s = sim_of(c);
  if (!s) return c ? fail(c, GMI_ERR_STATE, "no sim") : GMI_ERR_ARG;
  if (!name) return fail(c, GMI_ERR_ARG, "no name");
  HostState h;
  int rc;
  if ((rc = download(c, s, h))) return rc;
  const std::string nm(name);
  std::string text;
  if (read_file(nm + ".backup", &text)) return fail(c, GMI_ERR_ARG, ("cannot read " + nm + ".backup").c_str());
  Reader r(text);
  const double ntb = r.num();
  backup_records(s, h, [&](double &v) { v = r.num(); });
  if (!r.ok || !r.at_end() || ntb != std::floor(ntb))
    return fail(c, GMI_ERR_ARG, "backup file does not match this sim's sizes");
  // the GMRES warm start stays as it is: after make_body_reference_matrices,
  // the last reference solve's solution, as in a resumed reference run
  int is_exact = 0;
  std::string ext;
  if (read_file(nm + ".backup_gmi", &ext) == 0) {
    Reader e(ext);
    const std::string magic(kExtMagic);
    if (ext.compare(0, magic.size(), magic) != 0)
      return fail(c, GMI_ERR_ARG, "extended backup: bad header");
    e.p += magic.size();
    const double hdr[6] = {e.num(), e.num(), e.num(), e.num(), e.num(), e.num()};
    std::vector<float> lags;
    if ((rc = learn_lags(c, s, &lags, nullptr))) return rc;
    if (!e.ok || hdr[0] != ntb || hdr[1] != s->d.Ns || hdr[2] != s->d.Nf || hdr[3] != s->d.Nsw ||
        hdr[4] != s->d.npad)
      return fail(c, GMI_ERR_ARG, "extended backup does not match the backup or this sim");
    backup_records(s, h, [&](double &v) { v = e.num(); });
    const SimDev &d = s->d;
    const size_t ls = size_t(d.lds);
    for (int k = 0; k < 6; k++)
      for (int g = 0; g < d.Ns; g++) h.Vs[k * ls + g] = e.num();
    for (double &v : h.x) v = e.num();
    const double avg = e.num();
    std::vector<float> file_lags(static_cast<size_t>(hdr[5]));
    for (float &v : file_lags) v = float(e.num());
    if (!e.ok || !e.at_end()) return fail(c, GMI_ERR_ARG, "extended backup truncated");
    s->iters_avg = avg;
    if (!lags.empty() && !file_lags.empty()) {
      if (file_lags.size() != lags.size())
        return fail(c, GMI_ERR_ARG, "extended backup: lag features of another size");
      if ((rc = learn_lags(c, s, nullptr, &file_lags))) return rc;
    }
    is_exact = 1;
  }
  if ((rc = upload(c, s, h))) return rc;
  if (nt_of_backup) *nt_of_backup = int(ntb);
  if (exact) *exact = is_exact;
  return hip_check(c, hipDeviceSynchronize(), "sync");
}

}  // extern "C"
