"""Text formats of the reference's output and checkpoint files.

TEST INFRASTRUCTURE ONLY: the library writes these files itself
(csrc/gmi_io.hip); this module restates the reference's writers and reader
so that tests can compare the library's files byte for byte.

* Values in the .dat and .backup files are written through a stream set to
  std::scientific with precision 10, each followed by one space
  (cilia_sim_main.cpp:349-350,372, mobility_solver.cpp:1788-1811): the C
  format "%.10e ".
* The .par file and the reference-position files use the stream's default
  format (precision 6): "%g" (cilia_sim_main.cpp:81-93, rigid_body.cpp:391,
  swimmer.cpp:893).
"""
import numpy as np


def sci(v):
    return "%.10e " % v


def sci_line(prefix, values):
    return f"{prefix} " + "".join(sci(v) for v in np.ravel(values)) + "\n"


def backup_size(nfil, nseg):
    """Doubles in a backup, as the reference's reader counts them
    (cilia_sim_main.cpp:37-57, GMRES build, not prescribed cilia):
    per filament 9 + 28 NSEG, per swimmer 19 + its filaments."""
    return int(sum(19 + n * (9 + 28 * nseg) for n in nfil))


def backup_text(nt, values):
    """<name>.backup after step nt (cilia_sim_main.cpp:370-381)."""
    return sci_line(nt + 1, values)


def parse_backup(text):
    """(nt_of_backup, values) as the reader takes them
    (cilia_sim_main.cpp:61-67): the step count, then the doubles."""
    tok = text.split()
    return int(tok[0]), np.array([float(t) for t in tok[1:]])


def seg_states_line(nt, seg_q):
    """One line of <name>_seg_states.dat: nt+1, then every segment's
    quaternion (scalar part first), swimmer-major (filament.cpp:1577-1591)."""
    return sci_line(nt + 1, seg_q)


def body_states_line(nt, body_x, body_q):
    """One line of <name>_body_states.dat: nt+1, then per swimmer x (3) and
    q (4) (rigid_body.cpp:408-412)."""
    vals = np.concatenate([np.concatenate([x, q]) for x, q in zip(body_x, body_q)])
    return sci_line(nt + 1, vals)


def mobility_line(nt, values):
    """One line of _body_vels, _blob_forces, _seg_vels or _seg_forces: nt
    (not nt+1), then the values in storage order (mobility_solver.cpp:
    1783-1886)."""
    return sci_line(nt, values)


def par_text(nfil, nseg, nswim, nblob, mu, rseg, rblob, kb, kt, tol, total_steps, dt,
             plot_frequency):
    """<name>.par (cilia_sim_main.cpp:79-93, no cilia-type extras)."""
    rows = [(nfil, "NFIL"), (nseg, "NSEG"), (nswim, "NSWIM"), (nblob, "NBLOB"), (mu, "MU"),
            (rseg, "RSEG"), (rblob, "RBLOB"), (kb, "KB"), (kt, "KT"), (tol, "TOL"),
            (total_steps, "TOTAL_TIME_STEPS"), (dt, "DT"),
            (plot_frequency, "PLOT_FREQUENCY_IN_STEPS")]
    return "".join("%g %%%% %s\n" % (v, tag) for v, tag in rows)


def reference_positions_text(values):
    """<name>_blob_references.dat / _fil_references.dat (rigid_body.cpp:
    389-396, swimmer.cpp:889-897): "%g " per value, then a newline."""
    return "".join("%g " % v for v in np.ravel(values)) + "\n"
