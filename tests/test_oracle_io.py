"""The restated file formats (oracle/io.py) against the reference's own
machinery where it can run here, and the oracle's backup record against the
reference's reader arithmetic.

* "%.10e " and "%g" are checked against what a C++ std::ofstream prints with
  std::scientific + std::setprecision(10) and with its default settings --
  the stream configurations the reference uses (cilia_sim_main.cpp:349-350,
  81-93) -- compiled here with g++.
* The backup length of the oracle equals the reference reader's count
  (cilia_sim_main.cpp:37-57).
* A backup written and re-read through the text format restores an oracle
  run: the resumed steps match the uninterrupted ones to the Broyden
  tolerance (the text keeps 11 significant digits).
"""
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from generic_mobility_interface_class_amd import geometry
from oracle import binding as orc
from oracle import io

VALUES = [0.0, -0.0, 1.0, -1.0, 1.0 / 3.0, 1e-300, -2.5e300, 123456.789, 6.02214076e23,
          1e-5, 1e-4, 0.1, 1800.0, 1.0 / 300.0, 2.2, -7.25e-12, 4.0, 642.0]

CPP = r"""
#include <fstream>
#include <iomanip>
#include <iostream>
#include <cstdlib>
int main(int argc, char **argv) {
  std::ofstream a(argv[1]), b(argv[2]);
  b << 7 << " ";
  b << std::scientific << std::setprecision(10);
  for (int i = 3; i < argc; i++) {
    const double v = std::strtod(argv[i], nullptr);
    a << v << " ";
    b << v << " ";
  }
  return 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_formats_equal_cpp_streams():
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "f.cpp"), os.path.join(d, "f")
        open(src, "w").write(CPP)
        subprocess.run(["g++", "-O0", "-o", exe, src], check=True)
        fa, fb = os.path.join(d, "a"), os.path.join(d, "b")
        subprocess.run([exe, fa, fb] + [repr(v) for v in VALUES], check=True)
        got_default = open(fa).read()
        got_sci = open(fb).read()
    assert got_default == io.reference_positions_text(VALUES)[:-1]
    assert got_sci == "7 " + "".join(io.sci(v) for v in VALUES)


def test_backup_length_matches_reference_reader():
    for name in ("step_962",):
        cfg = geometry.sim_workload(name, seed=1)
        o = orc.Sim(cfg)
        assert o.backup_values().size == io.backup_size([16], 20) == 16 * (9 + 28 * 20) + 19
    cfg = geometry.ensemble_workload("ensemble_6", seed=4)
    o = orc.Sim(cfg)
    nfil = [g.nfil for g in cfg.swimmers]
    assert o.backup_values().size == io.backup_size(nfil, cfg.nseg)


def test_text_backup_resumes_oracle_run():
    cfg = geometry.sim_workload("step_962", seed=1)
    a = orc.Sim(cfg)
    a.make_body_reference_matrices()
    for nt in range(3):
        assert a.step(nt)[0] == 0
    text = io.backup_text(2, a.backup_values())
    nt0, vals = io.parse_backup(text)
    assert nt0 == 3
    b = orc.Sim(cfg)
    b.make_body_reference_matrices()
    b.set_backup(vals)
    # the restored state is the saved one to the 11 digits of the format
    np.testing.assert_allclose(b.backup_values(), a.backup_values(), rtol=1e-10, atol=1e-300)
    DL = cfg.swimmers[0].DL
    for nt in range(nt0, nt0 + 2):
        ra, rb = a.step(nt), b.step(nt)
        assert ra[0] == 0 and rb[0] == 0
        dx = np.abs(a.state()["x"] - b.state()["x"]).max()
        assert dx <= DL * cfg.tol, (nt, dx)
