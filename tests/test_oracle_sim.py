"""Pins the oracle's implicit-step restatement (oracle/sim.cpp, oracle/lie.hpp)
against facts independent of its own code:

* Lie-group helpers against scipy's rotation vectors (Rodrigues formula),
  finite differences of the exponential map, and the inverse/transpose
  identities the reference states (quaternion.cpp:516-518);
* own LU inverse against numpy;
* the rigid-body solve: blob velocities from the independent pair oracle
  equal a rigid motion U + W x r, and the blob forces balance the body
  force and torque (mobility_solver.cpp:905-964 solved to the GMRES
  tolerance);
* the body mobility of the blob sphere against Stokes drag (resolution-
  dependent known answer, mobility_solver.cpp:1746-1757) and its isotropy;
* the approximate Jacobian's exact position/Lie-algebra block against finite
  differences of the robot-arm recurrence;
* a straight unforced swimmer stays at rest; a converged step meets the
  BDF2 residual bound; error paths (non-convergence) report as the
  reference's driver would.
"""
import math

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from generic_mobility_interface_class_amd import geometry
from oracle import binding as orc


def quat_to_scipy(q):
    return Rotation.from_quat([q[1], q[2], q[3], q[0]])


@pytest.mark.parametrize("theta", [0.0, 3e-7, 0.3, 2.0])
def test_lie_exp_is_rotation_vector(theta):
    rng = np.random.default_rng(1)
    axis = rng.normal(size=3); axis /= np.linalg.norm(axis)
    u = theta * axis
    q = orc.lie_exp(u)
    assert abs(np.linalg.norm(q) - 1) < 1e-15
    R = orc.lie_rotmat(q)
    np.testing.assert_allclose(R, Rotation.from_rotvec(u).as_matrix(), atol=1e-14)
    np.testing.assert_allclose(R, quat_to_scipy(q).as_matrix(), atol=1e-14)


def test_midpoint_is_half_rotation():
    rng = np.random.default_rng(2)
    q1 = orc.lie_exp(rng.normal(size=3))
    u = 0.7 * rng.normal(size=3) / 2
    q2 = quat_mul(orc.lie_exp(u), q1)
    qm = orc.lie_midpoint(q1, q2)
    ref = quat_mul(orc.lie_exp(0.5 * u), q1)
    np.testing.assert_allclose(orc.lie_rotmat(qm), orc.lie_rotmat(ref), atol=1e-13)


def quat_mul(p, q):
    return np.array([p[0] * q[0] - p[1:] @ q[1:],
                     *(p[0] * q[1:] + q[0] * p[1:] + np.cross(p[1:], q[1:]))])


@pytest.mark.parametrize("theta", [0.0, 4e-6, 0.2, 1.5])
def test_dexp_dexpinv_inverse_and_transpose(theta):
    rng = np.random.default_rng(3)
    u = theta * rng.normal(size=3) / np.sqrt(3)
    v, w = rng.normal(size=3), rng.normal(size=3)
    np.testing.assert_allclose(orc.dexp(u, orc.dexpinv(u, v)), v, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(orc.dexpinv(u, orc.dexp(u, v)), v, rtol=1e-12, atol=1e-14)
    assert abs(w @ orc.dexpinv(u, v) - orc.dexpinv_transpose(u, w) @ v) < 1e-12


@pytest.mark.parametrize("theta", [0.3, 1.2])
def test_dexp_is_the_left_trivialised_differential(theta):
    """exp(u + e v) = exp(e dexp(u, v)) exp(u) + O(e^2): the convention the
    reference's BDF2 Lie residual (dexpinv of omega) relies on."""
    rng = np.random.default_rng(4)
    u = theta * rng.normal(size=3) / np.sqrt(3)
    v = rng.normal(size=3)
    e = 1e-6
    Rp = Rotation.from_rotvec(u + e * v).as_matrix()
    Rm = Rotation.from_rotvec(u - e * v).as_matrix()
    R0 = Rotation.from_rotvec(u).as_matrix()
    W = (Rp - Rm) / (2 * e) @ R0.T          # skew matrix of the spatial rate
    w = np.array([W[2, 1], W[0, 2], W[1, 0]])
    np.testing.assert_allclose(orc.dexp(u, v), w, rtol=1e-7, atol=1e-9)


def test_lu_inverse_against_numpy():
    rng = np.random.default_rng(5)
    A = rng.normal(size=(120, 120))
    Ai = orc.invert(A)
    np.testing.assert_allclose(Ai @ A, np.eye(120), atol=1e-10)
    with pytest.raises(np.linalg.LinAlgError):
        orc.invert(np.zeros((4, 4)))


def small_cfg(**kw):
    return geometry.sim_workload("step_962", seed=1, **kw)


def test_rigid_solve_is_a_rigid_motion_and_balances_forces():
    cfg = small_cfg(barrier_factor=0.0)
    s = orc.Sim(cfg)
    s.make_body_reference_matrices()
    s.initial_guess(0)
    s.forces()
    s.compute_velocities()
    m, st = s.mobility(), s.state()
    g = cfg.swimmers[0]
    q = st["body"][0, 3:7]
    x = st["body"][0, :3]
    R = orc.lie_rotmat(q)
    r = g.blob_references() @ R.T
    xb = x + r
    lam = m["f_blobs"]
    fs = m["f_segs"]
    # blob velocities from the independent pair oracle
    vb = orc.Mbs(xb, st["x"], fs, g.a_seg, g.a_blob) + orc.Mbb(xb, lam, g.a_blob)
    U = m["v_bodies"][0]
    rigid = U[:3] + np.cross(U[3:], r)
    assert np.abs(vb - rigid).max() <= 1e-6 * np.abs(rigid).max()
    # K^T lambda equals minus the filaments' net force and torque on the body
    d = st["x"] - x
    F = -fs[:, :3].sum(0)
    T = -(fs[:, 3:] + np.cross(d, fs[:, :3])).sum(0)
    scale = np.abs(fs).max() * len(fs)
    np.testing.assert_allclose(lam.sum(0), F, rtol=0, atol=1e-8 * scale)
    np.testing.assert_allclose(np.cross(r, lam).sum(0), T, rtol=0, atol=1e-8 * scale * cfg.swimmers[0].R)
    # segment velocities: M_ss F_s + M_sb lambda
    vs = orc.Msb(st["x"], xb, lam, g.a_seg, g.a_blob, V=orc.Mss(st["x"], fs, g.a_seg))
    assert np.abs(m["v_segs"] - vs).max() <= 1e-12 * np.abs(vs).max()


def test_body_mobility_matches_stokes_drag_and_is_isotropic():
    cfg = small_cfg()
    s = orc.Sim(cfg)
    s.make_body_reference_matrices()
    m = s.mobility()
    N, KK = m["Nref"][0], m["KTKinv"][0]
    R = cfg.swimmers[0].R
    tt, rr = 1 / (6 * math.pi * R), 1 / (8 * math.pi * R ** 3)
    # resolution-dependent: the blob shell is a slightly larger sphere
    assert abs(N[0, 0] / tt - 1) < 0.06 and abs(N[3, 3] / rr - 1) < 0.12
    np.testing.assert_allclose(np.diag(N)[:3], N[0, 0], rtol=1e-6)
    np.testing.assert_allclose(np.diag(N)[3:], N[3, 3], rtol=1e-6)
    off = N - np.diag(np.diag(N))
    assert np.abs(off).max() < 1e-5 * N[0, 0]
    # (K^T K)^-1 (K^T K) = I with K from the blob references
    r = cfg.swimmers[0].blob_references()
    K = np.zeros((6, 6))
    K[:3, :3] = len(r) * np.eye(3)
    for ri in r:
        rc = np.array([[0, ri[2], -ri[1]], [-ri[2], 0, ri[0]], [ri[1], -ri[0], 0]])
        K[:3, 3:] += rc
        K[3:, :3] -= rc
        K[3:, 3:] -= rc @ rc
    np.testing.assert_allclose(KK @ K, np.eye(6), atol=1e-12)


def test_straight_unforced_swimmer_stays_at_rest():
    cfg = small_cfg(curvature_amp=0.0)
    s = orc.Sim(cfg)
    s.make_body_reference_matrices()
    x0 = s.state()["x"].copy()
    for nt in range(3):
        rc, bi, gi = s.step(nt)
        assert rc == 0 and bi == 0
    np.testing.assert_allclose(s.state()["x"], x0, atol=1e-12)


def test_converged_step_meets_residual_bound():
    cfg = small_cfg()
    s = orc.Sim(cfg)
    s.make_body_reference_matrices()
    DL = cfg.swimmers[0].DL
    S = cfg.nseg
    for nt in range(3):
        rc, bi, gi = s.step(nt)
        assert rc == 0
        e, big = s.errors(nt)
        assert not big
        nf = cfg.swimmers[0].nfil
        fil = e[:6 * S * nf].reshape(nf, 2, S, 3)
        assert np.abs(fil[:, 0]).max() <= DL * cfg.tol
        assert np.abs(fil[:, 1]).max() <= cfg.tol
        assert np.abs(e[6 * S * nf:6 * S * nf + 3]).max() <= DL * cfg.tol


def test_jacobian_robot_arm_block_matches_finite_differences():
    """Rows = positions of filament 0, columns = its Lie-algebra unknowns:
    the reference's block (filament.cpp:1439-1455) is exact at u = 0."""
    cfg = small_cfg()
    s = orc.Sim(cfg)
    s.initial_guess(0)
    s.forces()
    J, Ji = s.filament_jacobian(0, 0)
    n = J.shape[0]
    np.testing.assert_allclose(J @ Ji, np.eye(n), atol=1e-9)
    S = cfg.nseg
    DL = cfg.swimmers[0].DL
    q = s.state()["q"][:S]

    def positions(qs):
        t = np.array([Rotation.from_quat([qq[1], qq[2], qq[3], qq[0]]).apply([1.0, 0, 0])
                      for qq in qs])
        x = np.zeros((S, 3))
        for i in range(1, S):
            x[i] = x[i - 1] + 0.5 * DL * (t[i - 1] + t[i])
        return x

    e = 1e-6
    for j in (1, 5, S - 1):
        for c in range(3):
            du = np.zeros(3); du[c] = e
            qp, qm = q.copy(), q.copy()
            qp[j] = quat_mul(orc.lie_exp(du), q[j])
            qm[j] = quat_mul(orc.lie_exp(-du), q[j])
            d = (positions(qp) - positions(qm)).reshape(-1) / (2 * e)
            col = J[:3 * S, 3 * (S + j) + c]
            np.testing.assert_allclose(col, d, atol=1e-7)


def test_non_convergence_is_reported():
    cfg = small_cfg(tol=1e-13, max_broyden_iter=1)
    s = orc.Sim(cfg)
    s.make_body_reference_matrices()
    rc, bi, gi = s.step(0)
    assert rc == -6 and bi == 1
