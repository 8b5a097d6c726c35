// gmi_drive.cpp -- command-line driver of the C++ mobility-backend class.
//
// Runs the velocity part of mobility_solver::compute_velocities
// (mobility_solver.cpp:1358-1402) in the reference's call order through
// gmi::rpy_mobility_solver, for given blob forces instead of a GMRES solve:
//   copy positions/forces -> barrier -> M_ss, M_bs -> copy blob velocities
//   -> (blob forces) -> M_sb -> copy segment velocities -> wait.
// Input file (raw little-endian): int32 ns, nb, nseg; double a_seg, a_blob,
// mu, K; then x_segs[3ns], f_segs[6ns], x_blobs[3nb], f_blobs[3nb].
// Output file: v_segs[6ns], v_blobs[3nb], f_blobs_repulsion[3nb].
#include <cstdio>
#include <vector>

#include "gmi_mobility_solver.hpp"

int main(int argc, char **argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: gmi_drive in.bin out.bin\n");
    return 2;
  }
  FILE *fi = std::fopen(argv[1], "rb");
  if (!fi) return 2;
  int hdr[3];
  double phys[4];
  if (std::fread(hdr, sizeof hdr, 1, fi) != 1 || std::fread(phys, sizeof phys, 1, fi) != 1) return 2;
  const int ns = hdr[0], nb = hdr[1], nseg = hdr[2];
  gmi_params p{};
  p.n_swimmers = 1;
  p.nseg_per_swimmer = &ns;
  p.nblob_per_swimmer = &nb;
  p.nseg_per_filament = nseg;
  p.swimmers_interact = 1;
  p.a_seg = phys[0];
  p.a_blob = phys[1];
  p.mu = phys[2];
  p.precision = GMI_FP64;
  p.device = 0;
  try {
    gmi::rpy_mobility_solver m(p);
    m.barrier_K = phys[3];
    m.initialise();
    std::vector<double> fb(3 * size_t(nb));
    if (std::fread(m.x_segs_host, sizeof(double), 3 * size_t(ns), fi) != 3 * size_t(ns) ||
        std::fread(m.f_segs_host, sizeof(double), 6 * size_t(ns), fi) != 6 * size_t(ns) ||
        std::fread(m.x_blobs_host, sizeof(double), 3 * size_t(nb), fi) != 3 * size_t(nb) ||
        std::fread(fb.data(), sizeof(double), fb.size(), fi) != fb.size())
      return 2;
    std::fclose(fi);
    m.copy_segment_positions_to_device();
    m.copy_segment_forces_to_device();
    m.copy_blob_positions_to_device();
    m.apply_interparticle_forces();
    m.copy_interparticle_blob_forces_to_host();
    m.wait_for_device();
    m.evaluate_segment_segment_mobility();
    m.evaluate_blob_segment_mobility();
    m.copy_blob_velocities_to_host();
    m.wait_for_device();
    for (size_t k = 0; k < fb.size(); k++) m.f_blobs_host[k] = fb[k];
    m.copy_blob_forces_to_device();
    m.evaluate_segment_blob_mobility();
    m.copy_segment_velocities_to_host();
    m.wait_for_device();
    FILE *fo = std::fopen(argv[2], "wb");
    if (!fo) return 2;
    std::fwrite(m.v_segs_host, sizeof(double), 6 * size_t(ns), fo);
    std::fwrite(m.v_blobs_host, sizeof(double), 3 * size_t(nb), fo);
    std::fwrite(m.f_blobs_repulsion_host, sizeof(double), 3 * size_t(nb), fo);
    std::fclose(fo);
    m.finalise();
  } catch (const gmi::gmi_error &e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  return 0;
}
