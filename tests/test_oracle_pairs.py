"""Pins the CPU oracle's pair formulas (oracle/rpy_pairs.cpp) against facts
that do not come from the oracle's own formula:

* self mobility closed forms 1/(6 pi mu a), 1/(8 pi mu a^3)
  (cuda_functions.cu:307,336-350,364-386 at r = 0);
* the far-field RPY tt tensor equals the double surface average of the
  Oseen tensor over both spheres (mean-value property of biharmonic
  functions) -- computed here by quadrature, so the a^2/3 corrections and
  the unequal-radius combination a_s^2 + a_b^2 are checked independently;
* the leading far-field terms are the Oseen, rotlet and potential-dipole
  tensors;
* far-field rotation couplings are half curls of the translational fields
  (finite differences), which pins the sign and the r-power of every tr/rr
  term;
* continuity of every block across r = a_i + a_j and r = |a_s - a_b|;
* symmetry and positive definiteness of the assembled grand mobility,
  which pins Msb against Mbs;
* the equal-radius reduction (unequal formulas at a_s = a_b equal the ss
  formulas);
* rotation/translation invariance;
* the barrier's hand-computed values and exclusions.
"""
import math

import numpy as np
import pytest

from oracle import binding as orc

MU = 1.0


def blocks_single_pair(xi, xj, a_t, a_s, kind):
    """Dense coupling of target i from source j through the oracle's O(N^2)
    apply (unit forces), as a (rows x cols) matrix.  kind: ss, sb, bs, bb."""
    xi = np.asarray(xi, float)[None]
    xj = np.asarray(xj, float)[None]
    if kind == "ss":
        X = np.vstack([xi, xj])
        M = np.zeros((6, 6))
        for q in range(6):
            F = np.zeros((2, 6)); F[1, q] = 1.0
            M[:, q] = orc.Mss(X, F, a_t, MU)[0]
        return M
    if kind == "bb":
        X = np.vstack([xi, xj])
        M = np.zeros((3, 3))
        for q in range(3):
            F = np.zeros((2, 3)); F[1, q] = 1.0
            M[:, q] = orc.Mbb(X, F, a_t, MU)[0]
        return M
    if kind == "sb":  # target seg (radius a_t), source blob (radius a_s)
        M = np.zeros((6, 3))
        for q in range(3):
            F = np.zeros((1, 3)); F[0, q] = 1.0
            M[:, q] = orc.Msb(xi, xj, F, a_t, a_s, MU)[0]
        return M
    if kind == "bs":  # target blob (a_t), source seg (a_s)
        M = np.zeros((3, 6))
        for q in range(6):
            F = np.zeros((1, 6)); F[0, q] = 1.0
            M[:, q] = orc.Mbs(xi, xj, F, a_s, a_t, MU)[0]
        return M
    raise ValueError(kind)


def test_self_mobility_closed_form():
    a = 1.3
    X = np.zeros((1, 3))
    F = np.array([[0.3, -1.2, 2.0, 0.7, 0.1, -0.4]])
    V = orc.Mss(X, F, a, MU)[0]
    np.testing.assert_allclose(V[:3], F[0, :3] / (6 * math.pi * MU * a), rtol=1e-15)
    np.testing.assert_allclose(V[3:], F[0, 3:] / (8 * math.pi * MU * a ** 3), rtol=1e-15)
    Vb = orc.Mbb(X, F[:, :3], a, MU)[0]
    np.testing.assert_allclose(Vb, F[0, :3] / (6 * math.pi * MU * a), rtol=1e-15)
    # unequal radii, coincident centres: translational drag of the larger
    # sphere, no rotation coupling from the smaller one
    M = blocks_single_pair([0, 0, 0], [0, 0, 0], 1.0, 1.4, "sb")
    np.testing.assert_allclose(M[:3], np.eye(3) / (6 * math.pi * MU * 1.4), rtol=1e-15)
    np.testing.assert_allclose(M[3:], 0.0, atol=1e-300)


def _sphere_quadrature(n):
    """Fibonacci points on the unit sphere with equal weights."""
    i = np.arange(n) + 0.5
    z = 1 - 2 * i / n
    phi = i * math.pi * (3 - math.sqrt(5))
    rho = np.sqrt(1 - z * z)
    return np.stack([rho * np.cos(phi), rho * np.sin(phi), z], 1)


def _oseen(r):
    d = np.linalg.norm(r, axis=-1)[..., None, None]
    rr = r[..., :, None] * r[..., None, :]
    return (np.eye(3) / d + rr / d ** 3) / (8 * math.pi * MU)


@pytest.mark.parametrize("a_t,a_s,dist", [(1.0, 1.0, 2.5), (1.0, 1.4, 3.0), (1.4, 1.0, 5.0)])
def test_far_tt_equals_surface_averaged_oseen(a_t, a_s, dist):
    n = 1500
    P = _sphere_quadrature(n)
    rng = np.random.default_rng(1)
    e = rng.normal(size=3); e /= np.linalg.norm(e)
    xi, xj = dist * e, np.zeros(3)
    # average over both surfaces: Oseen(xi + a_t p - xj - a_s q)
    acc = np.zeros((3, 3))
    for k in range(0, n, 250):
        r = xi + a_t * P[k:k + 250, None, :] - xj - a_s * P[None, :, :]
        acc += _oseen(r).sum(axis=(0, 1))
    avg = acc / (n * n)
    kind = "bb" if a_t == a_s else ("sb" if a_t < a_s else "bs")
    if kind == "bb":
        M = blocks_single_pair(xi, xj, a_t, a_s, "bb")
    elif kind == "sb":   # seg radius a_t (target), blob radius a_s (source)
        M = blocks_single_pair(xi, xj, a_t, a_s, "sb")[:3, :3]
    else:                # blob target radius a_t, seg source radius a_s
        M = blocks_single_pair(xi, xj, a_t, a_s, "bs")[:, :3]
    np.testing.assert_allclose(M, avg, rtol=0, atol=2e-4 * np.abs(avg).max())


def test_far_field_leading_terms():
    a = 1.0
    rng = np.random.default_rng(2)
    e = rng.normal(size=3); e /= np.linalg.norm(e)
    r = 1e4
    M = blocks_single_pair(r * e, np.zeros(3), a, a, "ss")
    oseen = (np.eye(3) + np.outer(e, e)) / (8 * math.pi * MU * r)
    dip = (3 * np.outer(e, e) - np.eye(3)) / (16 * math.pi * MU * r ** 3)
    rot = np.array([[0, e[2], -e[1]], [-e[2], 0, e[0]], [e[1], -e[0], 0]]) / (8 * math.pi * MU * r ** 2)
    np.testing.assert_allclose(M[:3, :3], oseen, rtol=0, atol=1e-7 * np.abs(oseen).max())
    np.testing.assert_allclose(M[3:, 3:], dip, rtol=1e-12, atol=0)
    # w = g f x r_hat  <=>  column q of the w<-f block is e_q x r_hat
    np.testing.assert_allclose(M[3:, :3], rot, rtol=1e-12, atol=1e-30)
    np.testing.assert_allclose(M[:3, 3:], rot, rtol=1e-12, atol=1e-30)


def _curl(field, x, h=1e-5):
    """Curl of a vector field R^3 -> R^3 by central differences."""
    J = np.zeros((3, 3))
    for k in range(3):
        dx = np.zeros(3); dx[k] = h
        J[:, k] = (field(x + dx) - field(x - dx)) / (2 * h)
    return np.array([J[2, 1] - J[1, 2], J[0, 2] - J[2, 0], J[1, 0] - J[0, 1]])


@pytest.mark.parametrize("a_t,a_s,kind", [(1.0, 1.0, "ss"), (1.0, 1.4, "sb")])
def test_far_rotation_is_half_curl(a_t, a_s, kind):
    rng = np.random.default_rng(3)
    f = rng.normal(size=3)
    x0 = np.array([2.9, -1.7, 2.2])

    def v_of(x):
        M = blocks_single_pair(x, np.zeros(3), a_t, a_s, kind)
        return M[:3, :3] @ f

    M = blocks_single_pair(x0, np.zeros(3), a_t, a_s, kind)
    w = M[3:, :3] @ f
    np.testing.assert_allclose(w, 0.5 * _curl(v_of, x0), rtol=1e-6, atol=1e-12)
    if kind == "ss":
        tau = rng.normal(size=3)

        def v_tau(x):
            return blocks_single_pair(x, np.zeros(3), a_t, a_s, "ss")[:3, 3:] @ tau

        np.testing.assert_allclose(M[3:, 3:] @ tau, 0.5 * _curl(v_tau, x0), rtol=1e-6,
                                   atol=1e-12)


@pytest.mark.parametrize("kind,a_t,a_s", [("ss", 1.0, 1.0), ("bb", 1.4, 1.4),
                                          ("sb", 1.0, 1.4), ("bs", 1.4, 1.0),
                                          ("sb", 1.4, 1.0)])
def test_continuity_across_branches(kind, a_t, a_s):
    e = np.array([0.36, -0.48, 0.8])
    radii = [a_t + a_s]
    if a_t != a_s:
        radii.append(abs(a_t - a_s))
    for rc in radii:
        lo = blocks_single_pair(rc * (1 - 1e-10) * e, np.zeros(3), a_t, a_s, kind)
        hi = blocks_single_pair(rc * (1 + 1e-10) * e, np.zeros(3), a_t, a_s, kind)
        np.testing.assert_allclose(lo, hi, rtol=0, atol=1e-8 * np.abs(hi).max())


def _grand_dense(xs, xb, a_s, a_b):
    ns, nb = len(xs), len(xb)
    n = 6 * ns + 3 * nb
    M = np.zeros((n, n))
    for q in range(n):
        fs = np.zeros((ns, 6)); fb = np.zeros((nb, 3))
        if q < 6 * ns:
            fs.reshape(-1)[q] = 1.0
        else:
            fb.reshape(-1)[q - 6 * ns] = 1.0
        vs, vb = orc.grand(xs, xb, fs, fb, a_s, a_b, MU, nthreads=1)
        M[:, q] = np.concatenate([vs.reshape(-1), vb.reshape(-1)])
    return M


@pytest.mark.parametrize("a_b", [1.0, 1.4])
def test_grand_mobility_symmetric_positive_definite(a_b):
    rng = np.random.default_rng(4)
    # a tight random cluster: many overlapping pairs in every branch
    xs = rng.uniform(-2.5, 2.5, size=(9, 3))
    xb = rng.uniform(-2.5, 2.5, size=(8, 3))
    # coincident centres (equal radii: nearly coincident, since two equal
    # spheres at one point make the mobility exactly singular)
    xb[0] = xs[0] + ([0.0, 0, 0] if a_b != 1.0 else [0.05, 0, 0])
    xb[1] = xs[1] + [0.2, 0, 0]        # r < |a_s - a_b| when a_b = 1.4
    M = _grand_dense(xs, xb, 1.0, a_b)
    np.testing.assert_allclose(M, M.T, rtol=0, atol=1e-14 * np.abs(M).max())
    lam = np.linalg.eigvalsh(0.5 * (M + M.T))
    assert lam.min() > 1e-9 * lam.max(), lam.min()


def test_equal_radius_reduction():
    rng = np.random.default_rng(5)
    xs = rng.uniform(-3, 3, size=(12, 3))
    xb = rng.uniform(-3, 3, size=(10, 3))
    fb = rng.normal(size=(10, 3))
    fs = rng.normal(size=(12, 6))
    a = 1.0
    # Msb with a_b = a_s equals Mss applied to blobs carrying zero torque
    X = np.vstack([xs, xb])
    F = np.zeros((22, 6)); F[12:, :3] = fb
    ref = orc.Mss(X, F, a, MU)[:12]
    np.testing.assert_allclose(orc.Msb(xs, xb, fb, a, a, MU), ref, rtol=1e-13, atol=1e-15)
    # Mbs with equal radii equals the v-rows of Mss for blob targets
    F2 = np.zeros((22, 6)); F2[:12] = fs
    ref2 = orc.Mss(X, F2, a, MU)[12:, :3]
    np.testing.assert_allclose(orc.Mbs(xb, xs, fs, a, a, MU), ref2, rtol=1e-13, atol=1e-15)


def test_rotation_translation_invariance():
    rng = np.random.default_rng(6)
    xs = rng.uniform(-4, 4, size=(20, 3))
    xb = rng.uniform(-4, 4, size=(15, 3))
    fs = rng.normal(size=(20, 6)); fb = rng.normal(size=(15, 3))
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    if np.linalg.det(q) < 0:
        q[:, 0] *= -1
    t = rng.normal(size=3) * 5
    vs, vb = orc.grand(xs, xb, fs, fb, 1.0, 1.4, MU)
    fs2 = np.hstack([fs[:, :3] @ q.T, fs[:, 3:] @ q.T])
    vs2, vb2 = orc.grand(xs @ q.T + t, xb @ q.T + t, fs2, fb @ q.T, 1.0, 1.4, MU)
    np.testing.assert_allclose(vs2[:, :3], vs[:, :3] @ q.T, rtol=1e-11, atol=1e-13)
    np.testing.assert_allclose(vs2[:, 3:], vs[:, 3:] @ q.T, rtol=1e-11, atol=1e-13)
    np.testing.assert_allclose(vb2, vb @ q.T, rtol=1e-11, atol=1e-13)


def test_barrier_hand_values_and_exclusions():
    K = 7.0
    # filament of 3 segments along x (spacing 2.2 > 1.1*2), one extra
    # segment of another filament at distance 2.1 from segment 0
    nseg = 3
    xs = np.array([[0, 0, 0], [2.2, 0, 0], [4.4, 0, 0],
                   [0, 2.1, 0], [0, 50, 0], [0, 52.2, 0]], float)
    xb = np.array([[0, -2.3, 0], [0, -4.5, 0]], float)
    fs = np.zeros((6, 6))
    f, rep = orc.barrier(xs, xb, fs, nseg, 1.0, 1.2, K)
    # seg0 <- seg3: d = 2.1, a_sum = 2, fac = 1 - 5*0.1 = 0.5
    # seg0 <- blob0: d = 2.3, a_sum = 2.2, fac = 1 - (10/2.2)*0.1
    fb0 = 1 - (10 / 2.2) * 0.1
    np.testing.assert_allclose(f[0, :3], [0, -K * 0.5 ** 4 + K * fb0 ** 4, 0], rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(f[3, :3], [0, K * 0.5 ** 4, 0], rtol=1e-14)
    # adjacent segments of one filament are excluded (d = 2.2 = 1.1*2 anyway)
    np.testing.assert_allclose(f[1:3, :3], 0.0, atol=0)
    # blob0 <- seg0: a_sum 2.2, d 2.3 ; blob0 <- blob1: d = 2.2 < 2.64, a_sum 2.4,
    # fac = min(1, 1 - (10/2.4)(2.2-2.4)) = 1
    np.testing.assert_allclose(rep[0], [0, -K * fb0 ** 4 + K * 1.0, 0], rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(rep[1], [0, -K * 1.0, 0], rtol=1e-14)
    assert np.all(f[:, 3:] == 0)
