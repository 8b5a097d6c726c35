// gmi_mobility_solver.hpp -- C++ host class with the reference's
// mobility-backend method set, implemented over the C ABI (include/gmi.h).
//
// Mirrors rpy_mobility_solver (rpy_mobility_solver.hpp:15-63) behind the
// abstract mobility_solver interface (mobility_solver.hpp:18-102):
//   * the same 15 member functions, with the same names and meaning;
//   * the same public raw host buffers (x_segs_host, f_segs_host, ...), which
//     model objects may alias and write into directly;
//   * copies and evaluations are queued on the device; host buffers filled
//     by copy_*_to_host are valid after wait_for_device().
// Unlike the reference, every C-ABI status is checked: a failure throws
// gmi_error carrying gmi_last_error() (the reference ignored CUDA errors).
// Sizes are run-time (gmi_params) instead of config.hpp macros.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/gmi.h"

namespace gmi {

class gmi_error : public std::runtime_error {
 public:
  gmi_error(int code, const std::string &msg)
      : std::runtime_error("gmi error " + std::to_string(code) + ": " + msg), code(code) {}
  int code;
};

class rpy_mobility_solver {
 public:
  // Host buffers (mobility_solver.hpp:54-60), AoS FP64 in the reference order.
  double *v_segs_host = nullptr;
  double *v_blobs_host = nullptr;
  double *x_segs_host = nullptr;
  double *x_blobs_host = nullptr;
  double *f_segs_host = nullptr;
  double *f_blobs_host = nullptr;
  double *f_blobs_repulsion_host = nullptr;

  // K = REPULSIVE_FORCE_FACTOR*END_FORCE_MAGNITUDE of the barrier kernel
  double barrier_K = 0.0;

  explicit rpy_mobility_solver(const gmi_params &p) : params_(p) {
    nseg_.assign(p.nseg_per_swimmer, p.nseg_per_swimmer + p.n_swimmers);
    nblob_.assign(p.nblob_per_swimmer, p.nblob_per_swimmer + p.n_swimmers);
    params_.nseg_per_swimmer = nseg_.data();
    params_.nblob_per_swimmer = nblob_.data();
    for (int n : nseg_) ns_ += n;
    for (int n : nblob_) nb_ += n;
  }
  ~rpy_mobility_solver() {
    free_device_memory();
    free_host_memory();
  }
  rpy_mobility_solver(const rpy_mobility_solver &) = delete;
  rpy_mobility_solver &operator=(const rpy_mobility_solver &) = delete;

  int num_segments() const { return ns_; }
  int num_blobs() const { return nb_; }
  gmi_ctx *context() { return ctx_; }

  // mobility_solver::initialise / finalise (mobility_solver.cpp:14-130)
  void initialise() {
    allocate_host_memory();
    allocate_device_memory();
    for (int n = 0; n < 3 * nb_; n++) f_blobs_host[n] = 0.0;
  }
  void finalise() {
    free_host_memory();
    free_device_memory();
  }

  // ---- the 15 backend members (mobility_solver.hpp:25-44) ----
  void allocate_host_memory() {
    host_.assign(size_t(15) * ns_ + size_t(12) * nb_, 0.0);
    double *p = host_.data();
    v_segs_host = p; p += 6 * ns_;
    f_segs_host = p; p += 6 * ns_;
    x_segs_host = p; p += 3 * ns_;
    v_blobs_host = p; p += 3 * nb_;
    x_blobs_host = p; p += 3 * nb_;
    f_blobs_host = p; p += 3 * nb_;
    f_blobs_repulsion_host = p;
  }
  void free_host_memory() {
    host_.clear();
    host_.shrink_to_fit();
    v_segs_host = v_blobs_host = x_segs_host = x_blobs_host = nullptr;
    f_segs_host = f_blobs_host = f_blobs_repulsion_host = nullptr;
  }
  void allocate_device_memory() {
    if (ctx_) return;
    gmi_ctx *c = nullptr;
    const int rc = gmi_create(&c, &params_);
    if (rc != GMI_OK) {
      const std::string msg = c ? gmi_last_error(c) : "invalid parameters";
      if (c) gmi_destroy(c);
      throw gmi_error(rc, msg);
    }
    ctx_ = c;
  }
  void free_device_memory() {
    if (ctx_) gmi_destroy(ctx_);
    ctx_ = nullptr;
  }

  void copy_segment_positions_to_device() { check(gmi_copy_segment_positions_to_device(ctx_, x_segs_host)); }
  void copy_segment_forces_to_device() { check(gmi_copy_segment_forces_to_device(ctx_, f_segs_host)); }
  void copy_blob_positions_to_device() { check(gmi_copy_blob_positions_to_device(ctx_, x_blobs_host)); }
  void copy_blob_forces_to_device() { check(gmi_copy_blob_forces_to_device(ctx_, f_blobs_host)); }

  void copy_interparticle_blob_forces_to_host() {
    check(gmi_copy_interparticle_blob_forces_to_host(ctx_, f_blobs_repulsion_host));
  }
  void copy_blob_velocities_to_host() { check(gmi_copy_blob_velocities_to_host(ctx_, v_blobs_host)); }
  void copy_segment_velocities_to_host() { check(gmi_copy_segment_velocities_to_host(ctx_, v_segs_host)); }

  void apply_interparticle_forces() { check(gmi_apply_interparticle_forces(ctx_, barrier_K)); }
  void wait_for_device() { check(gmi_wait_for_device(ctx_)); }
  void evaluate_segment_segment_mobility() { check(gmi_evaluate_segment_segment_mobility(ctx_)); }
  void evaluate_segment_blob_mobility() { check(gmi_evaluate_segment_blob_mobility(ctx_)); }
  void evaluate_blob_blob_mobility() { check(gmi_evaluate_blob_blob_mobility(ctx_)); }
  void evaluate_blob_segment_mobility() { check(gmi_evaluate_blob_segment_mobility(ctx_)); }

  // All four blocks in one launch (no reference counterpart).
  void evaluate_grand_mobility() { check(gmi_evaluate_grand_mobility(ctx_)); }

 private:
  void check(int rc) {
    if (rc != GMI_OK) throw gmi_error(rc, ctx_ ? gmi_last_error(ctx_) : "no context");
  }
  gmi_params params_;
  std::vector<int> nseg_, nblob_;
  int ns_ = 0, nb_ = 0;
  gmi_ctx *ctx_ = nullptr;
  std::vector<double> host_;
};

}  // namespace gmi
