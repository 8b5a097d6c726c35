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
  reference's driver would;
* the BDF2 / backward-Euler residual vanishes on paths the schemes
  integrate exactly (quadratic positions and Lie increments, angular
  velocities from scipy finite differences);
* internal forces: net force and torque of a free filament are those of the
  end force and perturbations alone, forces and torques rotate with the
  filament, a uniform bend or twist gets the closed-form restoring moment
  K 4 sin(theta/4)/DL, and the torques are minus the gradient of the
  discrete elastic energy to second order;
* the end force's approximate-Jacobian block against finite differences of
  the force;
* the Broyden secant condition H_{k+1} (e_{k+1} - e_k) = dx_k.
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


# ---------------------------------------------------------------------------
# Step mechanics pinned by exact solutions, balance laws and closed forms
# (compute_errors, internal forces and moments, the end force, the Broyden
# compact inverse update)

def _rotmats(qs):
    return Rotation.from_quat(np.column_stack([qs[:, 1], qs[:, 2], qs[:, 3], qs[:, 0]])).as_matrix()


def _quat(R):
    x, y, z, w = Rotation.from_matrix(R).as_quat()
    return np.array([w, x, y, z])


def _quadratic_path(rng, dt):
    """x(tau) = a + b tau + c tau^2 sampled at tau = -dt, 0, dt (tau measured
    from t_{n-1}): returns x, xm1, xm2 and v = x'(dt)."""
    a, b, c = rng.normal(size=3), rng.normal(size=3), rng.normal(size=3)
    x = lambda t: a + b * t + c * t * t
    return x(dt), x(0.0), x(-dt), b + 2 * c * dt


def _lie_path(rng, dt, quadratic=True):
    """Rotation path q(tau) = exp(u(tau)) q_{n-1} with u(tau) = al tau +
    be tau^2 (be = 0: linear): u = u(dt), um1 = -u(-dt) (the increment from
    q_{n-2} to q_{n-1}) and the spatial angular velocity at tau = dt from a
    central finite difference of the rotation matrices (scipy)."""
    al, be = rng.normal(size=3), (rng.normal(size=3) if quadratic else np.zeros(3))
    u = lambda t: al * t + be * t * t
    R = lambda t: Rotation.from_rotvec(u(t)).as_matrix()
    h = 1e-6
    W = (R(dt + h) - R(dt - h)) / (2 * h) @ R(dt).T
    w = np.array([W[2, 1], W[0, 2], W[1, 0]])
    return u(dt), -u(-dt), w


@pytest.mark.parametrize("nt", [0, 5])
def test_bdf2_residual_vanishes_on_exact_paths(nt):
    """compute_errors (mobility_solver.cpp:1404-1504,1641-1689) returns zero
    when every segment and the body follow paths the scheme integrates
    exactly: quadratic positions and a quadratic Lie-algebra increment for
    BDF2 (nt >= NUM_EULER_STEPS), linear ones for the backward-Euler step
    (nt = 0).  Velocities and angular velocities come from the paths
    themselves (finite differences of scipy rotation matrices), not from the
    oracle."""
    cfg = geometry.sim_config(nfil=3, nseg=6, nblob=42, R=4.0, seed=2)
    s = orc.Sim(cfg)
    dt = cfg.dt * 7.0          # a larger step makes the check sharper
    cfg_dt = cfg.dt
    rng = np.random.default_rng(11 + nt)
    bdf2 = nt >= cfg.num_euler_steps
    seg, vseg = np.zeros((s.Ns, 15)), np.zeros((s.Ns, 6))
    body, vbody = np.zeros((1, 15)), np.zeros((1, 6))

    def sample():
        x, xm1, xm2, v = _quadratic_path(rng, dt)
        if not bdf2:   # linear path through xm1 with velocity v
            x = xm1 + v * dt
        u, um1, w = _lie_path(rng, dt, quadratic=bdf2)
        return np.concatenate([x, xm1, xm2, u, um1]), v, w

    for g in range(s.Ns):
        row, v, w = sample()
        seg[g], vseg[g] = row, np.concatenate([v, w])
    row, v, w = sample()
    body[0], vbody[0] = row, np.concatenate([v, w])
    # the residual's time step is the configuration's: rescale the paths
    scale = cfg_dt / dt
    for arr in (vseg, vbody):
        arr *= 1.0 / scale
    s.set_kinematics(seg, vseg, body, vbody)
    e, big = s.errors(nt)
    S, nf = cfg.nseg, cfg.swimmers[0].nfil
    mag = np.abs(seg).max()
    assert not big
    assert np.abs(e).max() <= 1e-9 * mag, np.abs(e).max()
    # a wrong velocity on one segment shows up in that segment's rows only
    vseg[3, 1] += 1.0
    s.set_kinematics(seg, vseg, body, vbody)
    e2, big2 = s.errors(nt)
    coef = (2.0 / 3.0 if bdf2 else 1.0) * cfg_dt
    fil, j = divmod(3, S)
    row = 6 * S * fil + 3 * j + 1
    assert abs(e2[row] + coef) <= 1e-9 and big2


def _straight_then_bent(S, rng, bend=0.2):
    """Unit quaternions of a randomly bent and twisted filament."""
    R0 = Rotation.random(random_state=rng.integers(1 << 30)).as_matrix()
    qs = [_quat(R0)]
    R = R0
    for _ in range(S - 1):
        R = Rotation.from_rotvec(bend * rng.normal(size=3)).as_matrix() @ R
        q = _quat(R)
        qs.append(q if q @ qs[-1] >= 0 else -q)   # a continuous sign along the filament
    return np.array(qs)


def _arm(qs, DL, x0=None):
    t = _rotmats(qs)[:, :, 0]
    x = np.zeros((len(qs), 3)) if x0 is None else np.tile(x0, (len(qs), 1)).astype(float)
    for i in range(1, len(qs)):
        x[i] = x[i - 1] + 0.5 * DL * (t[i - 1] + t[i])
    return x


@pytest.mark.parametrize("end_force", [0.0, 204.5])
def test_internal_forces_balance_and_rotate_with_the_filament(end_force):
    """filament.cpp:1112-1220 on a free filament (tether and clamp zero):
    the constraint forces are internal, so the net force is the end force
    plus the perturbations alone and the net torque about any point (forces
    at robot-arm positions) is the end force's moment alone -- whatever the
    multipliers, bends and twists.  A rigid rotation of the filament and its
    multipliers rotates every force and torque."""
    rng = np.random.default_rng(21)
    S, DL, kb, kt = 20, 2.2, 1800.0, 1300.0
    qs = _straight_then_bent(S, rng)
    lam = rng.normal(size=3 * (S - 1)) * 50
    pert = rng.normal(size=6)
    z3 = np.zeros(3)
    st = np.array([0.0, 0.01, -0.02])
    f = orc.filament_forces(qs, lam, z3, z3, z3, st, kb, kt, DL, end_force, pert)
    x = _arm(qs, DL, x0=rng.normal(size=3))
    t_end = _rotmats(qs[-1:])[0, :, 0]
    if end_force == 0.0:     # the perturbations belong to the end-force build
        pert = np.zeros(6)
    F_ext = -end_force * t_end + pert[:3] + pert[3:]
    np.testing.assert_allclose(f[:, :3].sum(0), F_ext, atol=1e-9 * np.abs(lam).max())
    for p in (np.zeros(3), rng.normal(size=3) * 10):
        T = (np.cross(x - p, f[:, :3]) + f[:, 3:]).sum(0)
        T_ext = (np.cross(x[-1] - p, -end_force * t_end) + np.cross(x[4] - p, pert[:3]) +
                 np.cross(x[12] - p, pert[3:]))
        np.testing.assert_allclose(T, T_ext, atol=1e-8 * np.abs(f).max() * S * DL)
    # covariance under a rigid rotation Q
    Q = Rotation.random(random_state=7).as_matrix()
    qQ = _quat(Q)
    qs2 = np.array([quat_mul(qQ, q) for q in qs])
    lam2 = (lam.reshape(-1, 3) @ Q.T).reshape(-1)
    pert2 = (pert.reshape(2, 3) @ Q.T).reshape(-1)
    f2 = orc.filament_forces(qs2, lam2, z3, z3, z3, st, kb, kt, DL, end_force, pert2)
    if end_force == 0.0:
        f3 = orc.filament_forces(qs, lam, z3, z3, z3, st, kb, kt, DL, 0.0, None)
        np.testing.assert_array_equal(f3, f)
    np.testing.assert_allclose(f2[:, :3], f[:, :3] @ Q.T, atol=1e-9 * np.abs(f).max())
    np.testing.assert_allclose(f2[:, 3:], f[:, 3:] @ Q.T, atol=1e-9 * np.abs(f).max())


@pytest.mark.parametrize("axis,modulus", [(2, "kb"), (1, "kb"), (0, "kt")])
def test_uniform_bend_gets_the_restoring_closed_form_moment(axis, modulus):
    """A filament bent (about its binormal or normal) or twisted (about its
    tangent) uniformly by theta per link: q_{i+1} = exp(theta a) q_i.  The
    discrete strain 2 conj(q_mid)(q_{i+1} - q_i)/DL is 4 sin(theta/4)/DL
    along a, so each link carries the moment M = K 4 sin(theta/4)/DL a
    (K = KB for bending, KT for twist, filament.cpp:1096-1110): +M on the
    first segment, -M on the last -- the distal end is turned back against
    the bend -- and nothing on the interior ones.  A preferred strain equal
    to it (strain_twist) leaves no moment at all."""
    S, DL, kb, kt, theta = 12, 2.2, 1800.0, 1300.0, 0.3
    a = np.eye(3)[axis]                       # frame axes of the identity segment
    K = kb if modulus == "kb" else kt
    qs = np.array([_quat(Rotation.from_rotvec(i * theta * a).as_matrix()) for i in range(S)])
    z3 = np.zeros(3)
    f = orc.filament_forces(qs, np.zeros(3 * (S - 1)), z3, z3, z3, z3, kb, kt, DL)
    M = K * 4 * math.sin(theta / 4) / DL * a
    np.testing.assert_allclose(f[:, :3], 0, atol=1e-12)
    np.testing.assert_allclose(f[0, 3:], M, rtol=1e-12, atol=1e-9)
    np.testing.assert_allclose(f[-1, 3:], -M, rtol=1e-12, atol=1e-9)
    np.testing.assert_allclose(f[1:-1, 3:], 0, atol=1e-9 * K)
    assert f[-1, 3:] @ a < 0                  # restoring on the distal end
    st = 4 * math.sin(theta / 4) / DL * a     # body frame of every segment: a
    f0 = orc.filament_forces(qs, np.zeros(3 * (S - 1)), z3, z3, z3, st, kb, kt, DL)
    np.testing.assert_allclose(f0, 0, atol=1e-9 * K)


def test_elastic_torques_are_the_bending_energy_gradient_to_second_order():
    """For small bends the link moments are minus the gradient of the
    discrete elastic energy E = sum_links DL/2 (k - k0)^T diag(KT,KB,KB)
    (k - k0), with k the body-frame rotation vector of R_i^T R_{i+1} divided
    by DL (computed here with scipy), under a spatial rotation of one
    segment: the mismatch falls like theta^2."""
    S, DL, kb, kt = 8, 2.2, 1800.0, 1300.0
    Kd = np.array([kt, kb, kb])
    z3 = np.zeros(3)

    def energy(qs):
        R = _rotmats(qs)
        E = 0.0
        for i in range(S - 1):
            k = Rotation.from_matrix(R[i].T @ R[i + 1]).as_rotvec() / DL
            E += 0.5 * DL * k @ (Kd * k)
        return E

    errs = []
    for bend in (2e-2, 2e-3):
        rng = np.random.default_rng(31)
        qs = _straight_then_bent(S, rng, bend=bend)
        f = orc.filament_forces(qs, np.zeros(3 * (S - 1)), z3, z3, z3, z3, kb, kt, DL)
        g = np.zeros((S, 3))
        h = 1e-4 * bend
        for j in range(S):
            for c in range(3):
                d = np.zeros(3); d[c] = h
                qp, qm = qs.copy(), qs.copy()
                Rj = _rotmats(qs[j:j + 1])[0]
                qp[j] = _quat(Rotation.from_rotvec(d).as_matrix() @ Rj)
                qm[j] = _quat(Rotation.from_rotvec(-d).as_matrix() @ Rj)
                g[j, c] = (energy(qp) - energy(qm)) / (2 * h)
        errs.append(np.abs(f[:, 3:] + g).max() / np.abs(g).max())
    assert errs[0] < 1e-2 and errs[1] < 1e-4 and errs[1] < errs[0] / 50, errs


def test_end_force_jacobian_block_matches_finite_differences():
    """The end-force entry of the approximate Jacobian (filament.cpp:
    1466-1470) is the self-mobility factor times the derivative of the tip
    force -END_FORCE t with respect to the tip segment's Lie algebra element,
    taken here by finite differences of the oracle's force function."""
    ef = 204.5
    cfg = geometry.sim_config(nfil=3, nseg=20, nblob=42, R=4.0, seed=2, end_force=ef)
    cfg0 = geometry.sim_config(nfil=3, nseg=20, nblob=42, R=4.0, seed=2)
    mats = []
    for c in (cfg, cfg0):
        s = orc.Sim(c)
        s.initial_guess(3)
        s.forces()
        mats.append(s.filament_jacobian(0, 3)[0])
        q = s.state()["q"][:c.nseg]
    S, DL = cfg.nseg, cfg.swimmers[0].DL
    blk = (mats[0] - mats[1])[3 * (S - 1):3 * S, 3 * (2 * S - 1):]
    z3 = np.zeros(3)
    fac = -cfg.dt / (9 * math.pi * cfg.swimmers[0].mu * cfg.swimmers[0].a_seg)   # BDF2 row
    fd = np.zeros((3, 3))
    e = 1e-6
    for c in range(3):
        du = np.zeros(3); du[c] = e
        qp, qm = q.copy(), q.copy()
        qp[-1] = quat_mul(orc.lie_exp(du), q[-1])
        qm[-1] = quat_mul(orc.lie_exp(-du), q[-1])
        fp = orc.filament_forces(qp, np.zeros(3 * (S - 1)), z3, z3, z3, z3, 1.0, 1.0, DL, ef)
        fm = orc.filament_forces(qm, np.zeros(3 * (S - 1)), z3, z3, z3, z3, 1.0, 1.0, DL, ef)
        fd[:, c] = (fp[-1, :3] - fm[-1, :3]) / (2 * e)
    np.testing.assert_allclose(blk, fac * fd, rtol=1e-6, atol=1e-9 * abs(fac) * ef)


def test_broyden_update_satisfies_the_secant_condition():
    """After end_of_iter the inverse-Jacobian approximation
    H_{k+1} = gamma J~^-1 + sum_i C_i D_i^T maps the residual change of the
    iteration onto the update it applied: H_{k+1} (e_{k+1} - e_k) = dx_k
    (broyden_solver.cpp:70-130), at the first and the second iteration (the
    second also exercises the earlier rank-one pairs); dx_k itself is
    -H_k e_k (broyden_solver.cpp:26-68).  H is assembled here from the
    stored columns and the oracle's J~^-1 products."""
    cfg = geometry.sim_config(nfil=4, nseg=10, nblob=92, R=4.0, seed=3)
    s = orc.Sim(cfg)
    s.make_body_reference_matrices()
    assert s.broyden_begin(0)
    g = cfg.jacobian_confidence

    def H(v, it):
        Cm, Dm, _ = s.broyden_columns(cfg.max_broyden_iter)
        out = g * s.jacobian_inv_mult(v)
        for i in range(it):
            out += Cm[:, i] * (Dm[:, i] @ v)
        return out

    for k in range(2):
        Hk_e = None
        dx, eo, en, big = s.broyden_iterate(0)
        _, _, it = s.broyden_columns(cfg.max_broyden_iter)
        assert it == k + 1
        np.testing.assert_allclose(-H(eo, k), dx, rtol=0, atol=1e-11 * np.abs(dx).max())
        lhs = H(en - eo, k + 1)
        np.testing.assert_allclose(lhs, dx, rtol=0, atol=1e-10 * np.abs(dx).max())


def test_end_force_swimmer_steps_and_perturbs_only_the_first_step():
    """The default instability-driven physics on a small swimmer: the
    perturbation forces enter the first step only (forces of the same state
    with and without the perturbation input differ by about their magnitude
    at nt = 0 and agree bit for bit afterwards), and steps converge."""
    g = geometry.SwimmerGeometry(nfil=4, nseg=20, nblob=162, R=6.0)
    cfg = geometry.instability_config([g], seed=5)
    cfg_np = geometry.instability_config([g], seed=5, perturbation=None)
    o, o2 = orc.Sim(cfg), orc.Sim(cfg_np)
    for s in (o, o2):
        s.make_body_reference_matrices()
    for nt in range(3):
        for s in (o, o2):
            s.initial_guess(nt)
            s.forces()
        d = np.abs(o.mobility()["f_segs"] - o2.mobility()["f_segs"]).max()
        if nt == 0:
            assert 0.5 * 0.01 * cfg.end_force < d <= 2 * 0.01 * cfg.end_force
        else:
            assert d == 0.0
    s = orc.Sim(cfg)
    s.make_body_reference_matrices()
    for nt in range(4):
        rc, bi, _ = s.step(nt)
        assert rc == 0 and bi >= 1


def coupled_pair_cfg(interact):
    """Two swimmers 60 apart (body radius 4, filaments 22 long: tips at least
    8 apart), coupled or independent."""
    sw = [geometry.SwimmerGeometry(nfil=4, nseg=10, nblob=92, R=4.0,
                                   centre=np.array([0.0, 0.0, 0.0])),
          geometry.SwimmerGeometry(nfil=6, nseg=10, nblob=92, R=4.0,
                                   centre=np.array([60.0, 0.0, 0.0]))]
    rng = geometry.Xoshiro256ss(5)
    return geometry.SimConfig(swimmers=sw, phases=2 * np.pi * rng.uniform(10),
                              barrier_factor=0.0, interact=interact)


def _all_particle_velocities(cfg, s):
    """Blob and segment velocities of the rigid solve's forces from the pair
    oracle over every particle of every swimmer, with each swimmer's blob
    offsets from its body centre."""
    m, st = s.mobility(), s.state()
    xs, fs, lam = st["x"], m["f_segs"], m["f_blobs"]
    xb, r = [], []
    for n, g in enumerate(cfg.swimmers):
        R = orc.lie_rotmat(st["body"][n, 3:7])
        rn = g.blob_references() @ R.T
        r.append(rn)
        xb.append(st["body"][n, :3] + rn)
    xb_all = np.concatenate(xb)
    vb = orc.Mbs(xb_all, xs, fs, cfg.swimmers[0].a_seg, cfg.swimmers[0].a_blob) + \
        orc.Mbb(xb_all, lam, cfg.swimmers[0].a_blob)
    vs = orc.Msb(xs, xb_all, lam, cfg.swimmers[0].a_seg, cfg.swimmers[0].a_blob,
                 V=orc.Mss(xs, fs, cfg.swimmers[0].a_seg))
    return m, r, vb, vs


@pytest.mark.parametrize("interact", [True, False])
def test_coupled_swimmers_solve_against_every_particle(interact):
    """Coupled swimmers (the reference's NSWIM > 1 run): the rigid solve's
    blob velocities, evaluated by the pair oracle over every particle of both
    swimmers, are each body's rigid motion, and the segment velocities are
    the all-particle sums.  Independent swimmers fail the same check by the
    coupling they leave out (percent level at 60 apart), so the check can
    tell the two apart."""
    cfg = coupled_pair_cfg(interact)
    s = orc.Sim(cfg)
    s.make_body_reference_matrices()
    s.initial_guess(0)
    s.forces()
    s.compute_velocities()
    m, r, vb, vs = _all_particle_velocities(cfg, s)
    dev = []
    b0 = 0
    for n, g in enumerate(cfg.swimmers):
        U = m["v_bodies"][n]
        rigid = U[:3] + np.cross(U[3:], r[n])
        dev.append(np.abs(vb[b0:b0 + g.nblob] - rigid).max() / np.abs(rigid).max())
        b0 += g.nblob
    sdev = np.abs(m["v_segs"] - vs).max() / np.abs(vs).max()
    if interact:
        assert max(dev) <= 1e-6, dev
        assert sdev <= 1e-12, sdev
    else:
        assert max(dev) > 1e-4, dev
        assert sdev > 1e-6, sdev


def test_coupled_swimmers_step():
    cfg = coupled_pair_cfg(True)
    s = orc.Sim(cfg)
    s.make_body_reference_matrices()
    total = 0
    for nt in range(4):
        rc, bi, gi = s.step(nt)
        assert rc == 0
        total += bi
    assert total >= 1


def _gauss_jordan_no_pivoting(A):
    """In-place Gauss-Jordan inverse without row exchanges (numpy)."""
    A = np.array(A, dtype=np.float64)
    ratio = np.inf
    for k in range(len(A)):
        ratio = min(ratio, abs(A[k, k]) / np.abs(A[:, k]).max())
        inv = 1.0 / A[k, k]
        row = A[k].copy()
        row[k] = 1.0
        row *= inv
        col = A[:, k].copy()
        A -= np.outer(col, row)
        A[:, k] = -col * inv
        A[k] = row
    return A, ratio


@pytest.mark.parametrize("kind", ["curvature", "end_force"])
def test_filament_jacobians_invert_accurately_without_pivoting(kind):
    """The premise of the device's pivot-free filament inverse: on the
    approximate Jacobians of stepped filaments, elimination without row
    exchanges keeps every pivot within 1e-5 of its column's largest entry
    and leaves a residual |J J^-1 - I| within 50x of the oracle's pivoted LU
    inverse (both near 1e-14), and the fixed-vector check the device applies
    passes with a wide margin."""
    if kind == "curvature":
        cfg = small_cfg()
    else:
        g0 = geometry.SwimmerGeometry(nfil=6, nseg=20, nblob=162, R=6.0)
        cfg = geometry.instability_config([g0], seed=7)
    s = orc.Sim(cfg)
    s.make_body_reference_matrices()
    x = 1.0 + ((np.arange(6 * cfg.nseg) * 37) % 64) / 64.0
    for nt in range(4):
        assert s.step(nt)[0] == 0
        for f in (0, len(cfg.phases) - 1):
            J, Ji = s.filament_jacobian(f, nt + 1)
            A, ratio = _gauss_jordan_no_pivoting(J)
            I = np.eye(len(J))
            e_np, e_lu = np.abs(J @ A - I).max(), np.abs(J @ Ji - I).max()
            assert ratio > 1e-5, ratio
            assert e_np <= 50 * max(e_lu, 1e-15), (e_np, e_lu)
            assert np.abs(A @ (J @ x) - x).max() <= 2e-6 / 100
