"""Batched ensembles of uneven, hydrodynamically independent swimmers.

CPU checks: the generator's sizes follow its stated rules, and in the oracle
a swimmer stepped inside a batch follows the same trajectory as that swimmer
stepped alone.  The batch shares one GMRES and one Broyden solve (as the
reference does for all its swimmers, mobility_solver.cpp:970-1354,
broyden_solver.cpp:26-130), so the two agree to the solver tolerances, not
bit for bit.
"""
import math

import numpy as np

from generic_mobility_interface_class_amd import geometry
from oracle import binding as orc


def test_ensemble_sizes_follow_the_rules():
    cfg = geometry.ensemble_workload("ensemble_64", seed=4)
    assert len(cfg.swimmers) == 64
    again = geometry.ensemble_workload("ensemble_64", seed=4)
    assert [g.nfil for g in cfg.swimmers] == [g.nfil for g in again.swimmers]
    assert np.array_equal(cfg.phases, again.phases)
    for g in cfg.swimmers:
        assert 16 <= g.nfil <= 512 and g.nseg == 20
        f = math.isqrt((g.nblob - 2) // 10)
        assert 10 * f * f + 2 == g.nblob and 8 <= f <= 32
        assert len(g.blob_references()) == g.nblob
        assert abs(g.R - 25.0 * math.sqrt(g.nblob / 4002.0)) < 1e-12
    # the 16- and 256-filament swimmers land on the single-swimmer geometries
    # (642 blobs on R = 10 has 0.3% more blobs per area than 4002 on R = 25)
    small = geometry.ensemble_config(n_swimmers=1, nfil_min=16, nfil_max=16)
    assert small.swimmers[0].nblob == 642 and abs(small.swimmers[0].R - 10.0) < 0.02
    mid = geometry.ensemble_config(n_swimmers=1, nfil_min=256, nfil_max=256)
    assert (mid.swimmers[0].nblob, mid.swimmers[0].R) == (4002, 25.0)
    # no two swimmers come closer than their reach
    xs = np.array([g.centre[0] for g in cfg.swimmers])
    reach = max(g.base_radius() + g.nseg * g.DL for g in cfg.swimmers)
    assert np.diff(xs).min() > 2 * reach
    assert len(cfg.phases) == sum(g.nfil for g in cfg.swimmers)


def alone(cfg, k):
    f0 = sum(g.nfil for g in cfg.swimmers[:k])
    g = cfg.swimmers[k]
    return geometry.SimConfig(swimmers=[g], phases=cfg.phases[f0:f0 + g.nfil])


def test_swimmer_in_batch_follows_its_solo_trajectory():
    cfg = geometry.ensemble_workload("ensemble_6", seed=4)
    batch = orc.Sim(cfg)
    batch.make_body_reference_matrices()
    picks = (0, 3)
    solo = {k: orc.Sim(alone(cfg, k)) for k in picks}
    for s in solo.values():
        s.make_body_reference_matrices()
    seg0 = np.cumsum([0] + [g.nfil * g.nseg for g in cfg.swimmers])
    DL = cfg.swimmers[0].DL
    for nt in range(3):
        assert batch.step(nt)[0] == 0
        sb = batch.state()
        for k, s in solo.items():
            assert s.step(nt)[0] == 0
            ss = s.state()
            dx = np.abs(sb["x"][seg0[k]:seg0[k + 1]] - ss["x"]).max()
            assert dx <= 10 * DL * cfg.tol, (nt, k, dx)
            assert np.abs(sb["body"][k] - ss["body"][0]).max() <= 10 * DL * cfg.tol
