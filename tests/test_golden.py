"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py).

CPU: the oracle reproduces every fixture (regression guard) and the fixtures
carry the closed-form facts (self mobility, Msb = Mbs^T, SPD dense matrix).
GPU: the device path matches the fixtures to the parity bars.
"""
import math
import os

import numpy as np
import pytest

from generic_mobility_interface_class_amd import capi, geometry
from oracle import binding as orc

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(G, name), allow_pickle=False)


def test_apply_962_fixture_reproduced_by_oracle():
    f = load("apply_962.npz")
    g = geometry.workload_geometry("apply_962")
    xs, xb = g.positions()
    fs, fb = geometry.random_forces(g.ns, g.nblob, seed=int(f["seed"]))
    np.testing.assert_array_equal(xs, f["xs"])
    np.testing.assert_array_equal(fs, f["fs"])
    vs = orc.Msb(xs, xb, fb, 1.0, 1.0, V=orc.Mss(xs, fs, 1.0))
    np.testing.assert_allclose(vs, f["vs"], rtol=0, atol=1e-14 * np.abs(f["vs"]).max())


def test_branch_table_facts():
    f = load("branch_table.npz")
    kinds, r, B = f["kind"], f["r"], f["blocks"]
    at, as_ = f["a_t"], f["a_s"]
    for k in range(len(r)):
        if r[k] == 0.0 and kinds[k] in ("ss", "bb"):
            np.testing.assert_allclose(B[k][:3, :3], np.eye(3) / (6 * math.pi * at[k]), rtol=1e-15)
        if r[k] == 0.0 and kinds[k] == "ss":
            np.testing.assert_allclose(B[k][3:, 3:], np.eye(3) / (8 * math.pi), rtol=1e-15)
    # Msb(r) == Mbs(-r)^T for the translational and coupling rows: the
    # tables use the same direction, so compare sb at r with bs at r after
    # flipping the sign of the tr coupling (odd in r_hat)
    sb = {float(rr): B[k] for k, rr in enumerate(r) if kinds[k] == "sb"}
    bs = {float(rr): B[k] for k, rr in enumerate(r) if kinds[k] == "bs"}
    for rr in sb:
        np.testing.assert_allclose(sb[rr][:3, :3], bs[rr][:3, :3].T, rtol=1e-14, atol=1e-300)
        np.testing.assert_allclose(sb[rr][3:, :3], -bs[rr][:3, 3:].T, rtol=1e-14, atol=1e-300)


def test_dense16_is_spd_and_reproduced():
    f = load("dense16.npz")
    M = f["M"]
    np.testing.assert_allclose(M, M.T, atol=1e-14 * np.abs(M).max())
    assert np.linalg.eigvalsh(M).min() > 0
    fs = np.zeros((8, 6)); fb = np.zeros((8, 3))
    fs[3, 4] = 1.0
    vs, vb = orc.grand(f["xs"], f["xb"], fs, fb, float(f["a_s"]), float(f["a_b"]), nthreads=1)
    np.testing.assert_allclose(np.concatenate([vs.reshape(-1), vb.reshape(-1)]), M[:, 6 * 3 + 4],
                               rtol=0, atol=1e-15)


def test_step_962_fixture_reproduced_by_oracle():
    f = load("step_962.npz")
    cfg = geometry.sim_workload("step_962", seed=1)
    np.testing.assert_array_equal(cfg.phases, f["phases"])
    s = orc.Sim(cfg)
    s.make_body_reference_matrices()
    for nt in range(3):
        rc, bi, gi = s.step(nt)
        assert rc == 0 and (bi, gi) == tuple(f["iters"][nt])
        np.testing.assert_allclose(s.state()["x"], f[f"x_{nt}"], rtol=0, atol=1e-12)


@pytest.mark.gpu
def test_gpu_apply_matches_fixture(gpu):
    f = load("apply_962.npz")
    ctx = capi.Context([len(f["xs"])], [len(f["xb"])], 20)
    ctx.upload(f["xs"], f["fs"], f["xb"], f["fb"])
    ctx.call("gmi_evaluate_grand_mobility")
    vs, vb = ctx.velocities()
    assert np.abs(vs - f["vs"]).max() <= 1e-12 * np.abs(f["vs"]).max()
    assert np.abs(vb - f["vb"]).max() <= 1e-12 * np.abs(f["vb"]).max()


@pytest.mark.gpu
def test_gpu_steps_match_fixture(gpu):
    f = load("step_962.npz")
    cfg = geometry.sim_workload("step_962", seed=1)
    sim = capi.Sim(cfg)
    sim.make_body_reference_matrices()
    DL = cfg.swimmers[0].DL
    for nt in range(3):
        rc, bi, gi = sim.step(nt)
        assert rc == 0 and bi == f["iters"][nt][0]
        st = sim.state()
        assert np.abs(st["x"] - f[f"x_{nt}"]).max() <= DL * cfg.tol
        assert np.abs(st["u"] - f[f"u_{nt}"]).max() <= cfg.tol
