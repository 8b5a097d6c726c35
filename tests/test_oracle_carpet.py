"""The prescribed-cilia oracle above a no-slip wall (oracle/cilia.cpp with
wall = 1, the reference's INFINITE_PLANE_WALL build) pinned on CPU.

* Seeding (swimmer.cpp:68-165): grid counts, spacings, centring and base
  heights follow the closed forms of the rectangular and hexagonal lattices.
* Geometry: every filament beats in the y-z plane through its base (the frame
  maps the beat's x axis to +z, swimmer.cpp:91, filament.cpp:245-282), stays
  above the wall, and starts from phase 2 pi x / L + 2 pi 116/300
  (filament.cpp:222).
* Operator: equals the saddle-point matrix assembled densely here from the
  wall mobility (orc_Mss_wall, pinned in test_oracle_wall.py) and the
  velocity directions; no blob or body rows.
* Preconditioner: the exact inverse of that matrix with the mobility replaced
  by I / (6 pi mu a).
* Solve: the GMRES solution has a true residual within the tolerance of the
  dense system and matches the dense direct solve.
* Far above the wall the operator tends to the unbounded one: lifting the
  carpet by h changes it by O(1/h) relative.
* Phase dynamics: a dynamic-phase filament driven by the reference table of
  its own configuration recovers phase speed 2 pi.
"""
import dataclasses
import math

import numpy as np
import pytest

from generic_mobility_interface_class_amd import geometry
from oracle import binding as orc


def test_rectangular_seeding():
    nfil, nseg = 12, 10
    refs = geometry.carpet_seed(nfil, nseg)
    L = 2.2 * (nseg - 1)
    nx = int(math.sqrt(nfil))
    ny = math.ceil(nfil / nx)
    assert (nx, ny) == (3, 4)
    assert refs.shape == (nfil, 3)
    np.testing.assert_allclose(refs[:, 2], 1.1)
    # fil_id = j + i ny: x steps with i, y with j, 1.4 L apart, centred
    for fid in range(nfil):
        i, j = divmod(fid, ny)
        assert refs[fid, 0] == pytest.approx((i - 1.0) * 1.4 * L)
        assert refs[fid, 1] == pytest.approx((j - 1.5) * 1.4 * L)
    assert abs(refs[:, 0].mean()) < 1e-12 and abs(refs[:, 1].mean()) < 1e-12


@pytest.mark.parametrize("nfil", [7, 20, 64])
def test_hexagonal_seeding(nfil):
    nseg = 20
    refs = geometry.carpet_seed(nfil, nseg, seeding="hexagonal")
    L = 2.2 * (nseg - 1)
    dy = 0.5 * (1.0122 * L + 2.2)
    dx = 2.0 * dy / math.sqrt(3.0)
    assert refs.shape == (nfil, 3)
    np.testing.assert_allclose(refs[:, 2], 1.1)
    # rows dy apart; within a row, neighbours dx apart; odd rows shifted dx/2
    rows = np.round((refs[:, 1] - refs[:, 1].min()) / dy, 9)
    assert np.allclose(rows, np.round(rows))
    for r in np.unique(rows):
        xs = np.sort(refs[rows == r, 0])
        if len(xs) > 1:
            np.testing.assert_allclose(np.diff(xs), dx)
    # nearest neighbours in a hexagonal lattice: all at distance dx (dy = dx
    # sqrt(3)/2, so the diagonal neighbour is also dx away)
    d = np.linalg.norm(refs[:, None, :2] - refs[None, :, :2], axis=2) + np.eye(nfil) * 1e9
    np.testing.assert_allclose(d.min(axis=1), dx, rtol=1e-12)


def small_carpet(**kw):
    return geometry.carpet_config(nfil=6, nseg=8, **kw)


@pytest.mark.parametrize("beat", [geometry.BEAT_BUILD, geometry.BEAT_FULFORD_BLAKE,
                                  geometry.BEAT_CORAL_LARVAE])
def test_geometry_above_the_wall(beat):
    cfg = small_carpet(beat=beat)
    o = orc.Cilia(cfg)
    o.initial_guess(0)
    st = o.state()
    refs = np.asarray(cfg.fil_refs)
    L = 2.2 * (cfg.nseg - 1)
    np.testing.assert_allclose(st["phase"], 2 * np.pi * refs[:, 0] / L + 2 * np.pi * 116 / 300,
                               rtol=1e-14)
    x = st["x"].reshape(cfg.nf, cfg.nseg, 3)
    k = st["vel_dir"].reshape(cfg.nf, cfg.nseg, 3)
    np.testing.assert_allclose(x[:, 0], refs, atol=1e-13)
    # the beat plane is y-z through the base
    assert np.abs(x[:, :, 0] - refs[:, None, 0]).max() <= 1e-12 * L
    assert np.abs(k[:, :, 0]).max() <= 1e-12 * max(1.0, np.abs(k).max())
    assert x[:, :, 2].min() > 0
    # the filament points away from the wall at the base (the frame's x is +z)
    assert np.all(x[:, 1, 2] > x[:, 0, 2])


def dense_carpet(o, cfg, diag=False):
    st = o.state()
    xs = st["x"]
    ns = xs.shape[0]
    if diag:
        M = np.eye(3 * ns) / (6 * np.pi * cfg.mu * cfg.a_seg)
    else:
        M = np.zeros((3 * ns, 3 * ns))
        for j in range(3 * ns):
            f = np.zeros((ns, 6))
            f[j // 3, j % 3] = 1.0
            M[:, j] = orc.Mss_wall(xs, f, cfg.a_seg, cfg.mu)[:, :3].reshape(-1)
    n = cfg.system_size()
    A = np.zeros((n, n))
    A[:3 * ns, :3 * ns] = M
    if cfg.dynamic_phase:
        K = st["vel_dir"]
        S = cfg.nseg
        for f in range(cfg.nf):
            kf = K[f * S:(f + 1) * S].reshape(-1)
            A[3 * f * S:3 * (f + 1) * S, 3 * ns + f] -= kf
            A[3 * ns + f, 3 * f * S:3 * (f + 1) * S] -= kf
    return A


def with_table(cfg, n=24):
    return dataclasses.replace(cfg, dynamic_phase=True,
                               gen_force_refs=orc.cilia_reference_gen_forces(cfg, n))


@pytest.mark.parametrize("dyn", [False, True])
def test_operator_and_preconditioner_are_the_dense_blocks(dyn):
    cfg = small_carpet(seeding="hexagonal")
    if dyn:
        cfg = with_table(cfg)
    o = orc.Cilia(cfg)
    assert o.n == cfg.system_size() == 3 * cfg.ns + (cfg.nf if dyn else 0)
    o.initial_guess(0)
    A = dense_carpet(o, cfg)
    rng = np.random.default_rng(4)
    v = rng.standard_normal(o.n)
    np.testing.assert_allclose(o.apply_operator(v), A @ v, rtol=0,
                               atol=1e-12 * np.abs(A @ v).max())
    A0 = dense_carpet(o, cfg, diag=True)
    np.testing.assert_allclose(A0 @ o.apply_preconditioner(v), v, rtol=0,
                               atol=1e-10 * np.abs(v).max())
    # the wall mobility block is symmetric positive definite
    M = A[:3 * cfg.ns, :3 * cfg.ns]
    np.testing.assert_allclose(M, M.T, atol=1e-14 * np.abs(M).max())
    assert np.linalg.eigvalsh(0.5 * (M + M.T)).min() > 0
    # rhs: prescribed phases give omega k; dynamic phases -q
    b = o.rhs()
    if dyn:
        assert np.abs(b[:3 * cfg.ns]).max() == 0.0
        ph = o.state()["phase"]
        np.testing.assert_allclose(b[3 * cfg.ns:], [-o.interp_gen_force(p) for p in ph],
                                   rtol=1e-14)
    else:
        np.testing.assert_allclose(b, 2 * np.pi * o.state()["vel_dir"].reshape(-1), rtol=1e-14)


@pytest.mark.parametrize("dyn", [False, True])
def test_solution_is_the_dense_solve(dyn):
    cfg = small_carpet(beat=geometry.BEAT_FULFORD_BLAKE)
    if dyn:
        cfg = with_table(cfg)
    o = orc.Cilia(cfg)
    o.initial_guess(0)
    rc, its = o.compute_velocities(0)
    assert rc == 0 and 0 < its < cfg.gmres_max_iter
    A = dense_carpet(o, cfg)
    b = o.rhs()
    x = np.linalg.solve(A, b)
    sol = o.solution()
    got = sol["f_segs"].reshape(-1)
    if dyn:
        got = np.concatenate([got, o.state()["speed"]])
    assert np.linalg.norm(A @ got - b) <= 2 * cfg.gmres_tol * np.linalg.norm(b)
    assert np.linalg.norm(got - x) <= np.linalg.cond(A) * 2 * cfg.gmres_tol * np.linalg.norm(x)
    # the segment velocities of the solution are the prescribed ones
    if not dyn:
        np.testing.assert_allclose(o.segment_velocities(), 2 * np.pi * o.state()["vel_dir"],
                                   rtol=0, atol=1e-8 * np.abs(o.state()["vel_dir"]).max())


def rotne_prager(xs, f, a=1.0, mu=1.0):
    """Unbounded Rotne-Prager tensor in its far form at every separation
    (the form the wall kernel uses for all pairs, as the reference does)."""
    d = xs[:, None, :] - xs[None, :, :]
    r = np.linalg.norm(d, axis=2)
    np.fill_diagonal(r, 1.0)
    e = d / r[..., None]
    c1 = (1.0 + 2.0 * a * a / (3.0 * r * r)) / (8 * np.pi * mu * r)
    c2 = (1.0 - 2.0 * a * a / (r * r)) / (8 * np.pi * mu * r)
    np.fill_diagonal(c1, 0.0)
    np.fill_diagonal(c2, 0.0)
    v = np.einsum("ij,jk->ik", c1, f) + np.einsum("ij,ijk,ijl,jl->ik", c2, e, e, f)
    return v + f / (6 * np.pi * mu * a)


def test_far_above_the_wall_the_operator_is_unbounded():
    """The image terms decay like 1/h: lifting the carpet by h, the wall
    operator approaches the Rotne-Prager operator of free space."""
    base = small_carpet(beat=geometry.BEAT_BUILD)
    f = np.random.default_rng(8).standard_normal((base.ns, 3))
    errs = []
    for h in (1e3, 1e4):
        refs = np.asarray(base.fil_refs).copy()
        refs[:, 2] += h
        cfg = dataclasses.replace(base, fil_refs=refs)
        o = orc.Cilia(cfg)
        o.initial_guess(0)
        xs = o.state()["x"]
        free = rotne_prager(xs, f, cfg.a_seg, cfg.mu)
        wall = o.apply_operator(f.reshape(-1)).reshape(-1, 3)
        errs.append(np.abs(wall - free).max() / np.abs(free).max())
    # O(1/h): ten times higher, about ten times closer
    assert errs[0] < 0.01 and 9 < errs[0] / errs[1] < 11, errs


def test_dynamic_phase_recovers_the_reference_speed():
    steps = 30
    base = geometry.carpet_config(nfil=1, nseg=10, beat=geometry.BEAT_CORAL_LARVAE,
                                  phase=np.zeros(1), dt=1.0 / steps)
    table = orc.cilia_reference_gen_forces(base, steps)
    cfg = dataclasses.replace(base, gen_force_refs=table, dynamic_phase=True)
    o = orc.Cilia(cfg)
    for nt in range(4):
        rc, _ = o.step(nt)
        assert rc == 0
        assert o.state()["speed"][0] == pytest.approx(2 * np.pi, rel=1e-6)


def test_wall_refusals():
    cfg = small_carpet()
    for bad in (dict(fixed_body=False), dict(body_velocity=np.ones((1, 6))),
                dict(body_x=np.ones((1, 3)))):
        with pytest.raises(ValueError):
            orc.Cilia(dataclasses.replace(cfg, **bad))
