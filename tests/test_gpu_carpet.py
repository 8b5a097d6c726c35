"""Prescribed-cilia carpet above a no-slip wall on the GPU (gmi_set_wall +
gmi_cilia_*, k_pair_wall<true>) against the oracle (oracle/cilia.cpp with
wall = 1, pinned in test_oracle_carpet.py), through the C ABI.

* Building blocks on the same geometry (shapes, velocity directions, right-
  hand side, operator, preconditioner) agree to 1e-12 relative for both
  lattices, all three beats, prescribed and dynamic phases.
* Steps: positions and phases follow the oracle to 1e-9 relative; the device
  solution has a true residual (oracle operator, oracle rhs) within twice the
  GMRES tolerance; generalised forces and segment velocities agree to 1e-6.
* With prescribed phases the solved segment velocities are omega k.
* A dynamic-phase filament driven by the reference table of its own
  configuration recovers phase speed 2 pi; the 20,480-segment carpet steps
  with a true residual within the GMRES tolerance (1e-8 at the iteration
  cap, whose iterate the reference accepts).
* A zero right-hand side (zero frequency) gives the zero solution without
  an Arnoldi step.
* Refusals: mixed precision, a free body, and switching the wall off while
  a carpet is stepping.
"""
import dataclasses

import numpy as np
import pytest

from generic_mobility_interface_class_amd import capi, geometry
from oracle import binding as orc

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(1e-300, np.abs(np.asarray(b)).max())


VARIANTS = [
    # (seeding, beat, dynamic)
    ("rectangular", geometry.BEAT_BUILD, False),
    ("hexagonal", geometry.BEAT_CORAL_LARVAE, True),
    ("rectangular", geometry.BEAT_FULFORD_BLAKE, True),
    ("hexagonal", geometry.BEAT_FULFORD_BLAKE, False),
]


def variant_cfg(seeding, beat, dyn, nfil=9, nseg=12):
    cfg = geometry.carpet_config(nfil=nfil, nseg=nseg, seeding=seeding, beat=beat)
    if dyn:
        cfg = dataclasses.replace(cfg, dynamic_phase=True,
                                  gen_force_refs=orc.cilia_reference_gen_forces(cfg, 24))
    return cfg


@pytest.mark.parametrize("variant", VARIANTS)
def test_building_blocks_match_oracle(gpu, variant):
    cfg = variant_cfg(*variant)
    print("carpet variant", variant, flush=True)
    o, g = orc.Cilia(cfg), capi.Cilia(cfg)
    assert g.n == o.n == cfg.system_size()
    o.initial_guess(0)
    g.initial_guess(0)
    so, sg = o.state(), g.state()
    assert rel(sg["x"], so["x"]) <= 1e-12
    assert rel(sg["vel_dir"], so["vel_dir"]) <= 1e-12
    assert rel(sg["phase"], so["phase"]) <= 1e-15
    assert rel(g.rhs(), o.rhs()) <= 1e-12
    rng = np.random.default_rng(12)
    for _ in range(2):
        v = rng.standard_normal(o.n)
        assert rel(g.apply_operator(v), o.apply_operator(v)) <= 1e-12
        assert rel(g.apply_preconditioner(v), o.apply_preconditioner(v)) <= 1e-12


def solution_vector(cfg, g):
    x = g.solution()["f_segs"].reshape(-1)
    return np.concatenate([x, g.state()["speed"]]) if cfg.dynamic_phase else x


@pytest.mark.parametrize("variant", VARIANTS[:3])
def test_steps_match_oracle(gpu, variant):
    cfg = variant_cfg(*variant)
    print("carpet variant", variant, flush=True)
    o, g = orc.Cilia(cfg), capi.Cilia(cfg)
    for nt in range(3):
        o.initial_guess(nt)
        g.initial_guess(nt)
        ro, io = o.compute_velocities(nt)
        rg, ig = g.compute_velocities(nt)
        assert ro == 0 and rg == 0
        print(f"step {nt}: gmres oracle {io} gpu {ig}", flush=True)
        assert abs(io - ig) <= 2
        b = o.rhs()
        r = np.linalg.norm(o.apply_operator(solution_vector(cfg, g)) - b)
        assert r <= 2 * cfg.gmres_tol * np.linalg.norm(b), r / np.linalg.norm(b)
        assert rel(g.generalised_forces(), o.generalised_forces()) <= 1e-6
        vs = g.segment_velocities()
        assert rel(vs, o.segment_velocities()) <= 1e-6
        if not cfg.dynamic_phase:
            k = g.state()["vel_dir"]
            assert np.abs(vs - 2 * np.pi * k).max() <= 1e-7 * np.abs(k).max()
        o.integrate_bodies(); g.integrate_bodies()
        o.end_of_step(nt); g.end_of_step(nt)
        so, sg = o.state(), g.state()
        for key in ("x", "phase"):
            assert rel(sg[key], so[key]) <= 1e-9, key


@pytest.mark.parametrize("name", ["carpet_1280", "carpet_20480"])
def test_dynamic_phase_and_full_carpet(gpu, name):
    steps = 60
    w = dict(geometry.CARPET_WORKLOADS[name])
    base = geometry.carpet_config(**w, dt=1.0 / steps)
    # the reference table of one filament of this carpet (phase 0, speed
    # 2 pi), then that filament with dynamic phases: the closed form is 2 pi
    one = dataclasses.replace(base, nfil=np.array([1]),
                              fil_refs=np.asarray(base.fil_refs).reshape(-1, 3)[:1],
                              phase=np.zeros(1))
    table = orc.cilia_reference_gen_forces(one, steps)
    cfg = dataclasses.replace(one, dynamic_phase=True, gen_force_refs=table)
    g = capi.Cilia(cfg)
    for nt in range(4):
        rc, its = g.step(nt)
        assert rc == 0
        sp = g.state()["speed"][0]
        print(f"{name} one filament step {nt}: {its} iterations, speed {sp:.12f}", flush=True)
        assert sp == pytest.approx(2 * np.pi, rel=1e-6)
    full = dataclasses.replace(base, dynamic_phase=True, gen_force_refs=table)
    gf = capi.Cilia(full)
    for nt in range(2):
        gf.initial_guess(nt)
        rc, its = gf.compute_velocities(nt)
        assert rc == 0
        b = gf.rhs()
        r = np.linalg.norm(gf.apply_operator(solution_vector(full, gf)) - b)
        print(f"{name} full step {nt}: {its} iterations, residual {r / np.linalg.norm(b):.2e}",
              flush=True)
        bound = 2 * full.gmres_tol if its < full.gmres_max_iter else 1e-8
        assert r <= bound * np.linalg.norm(b)
        gf.integrate_bodies()
        gf.end_of_step(nt)
    assert np.all(np.isfinite(gf.state()["x"]))


def test_zero_right_hand_side_gives_zero_solution(gpu):
    """Prescribed phases at zero frequency: the right-hand side omega k is
    zero, and the GMRES start (decided on the device) returns x = 0 with no
    Arnoldi step; the same carpet at the usual frequency still matches the
    oracle."""
    base = variant_cfg("rectangular", geometry.BEAT_BUILD, False, nfil=4, nseg=8)
    g = capi.Cilia(dataclasses.replace(base, omega0=np.zeros(base.nf)))
    g.initial_guess(0)
    rc, its = g.compute_velocities(0)
    assert rc == 0 and its == 0
    assert np.abs(g.rhs()).max() == 0.0
    assert np.abs(g.solution()["f_segs"]).max() == 0.0
    o, h = orc.Cilia(base), capi.Cilia(base)
    for c in (o, h):
        c.initial_guess(0)
        assert c.compute_velocities(0)[0] == 0
    assert rel(h.solution()["f_segs"], o.solution()["f_segs"]) <= 1e-6


def test_carpet_refusals(gpu):
    cfg = variant_cfg("rectangular", geometry.BEAT_BUILD, False, nfil=4, nseg=8)
    with pytest.raises(RuntimeError):
        capi.Cilia(cfg, precision=capi.GMI_FP32_KAHAN)
    with pytest.raises(RuntimeError):
        capi.Cilia(dataclasses.replace(cfg, fixed_body=False))
    g = capi.Cilia(cfg)
    with pytest.raises(RuntimeError):
        g.set_wall(False)
    with pytest.raises(RuntimeError):
        g.call("gmi_cilia_set_precision", capi.GMI_FP32_KAHAN)
    rc, _ = g.step(0)
    assert rc == 0
