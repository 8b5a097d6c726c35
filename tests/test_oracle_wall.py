"""Pins of the wall-bounded segment mobility restatement (oracle
orc_Mss_wall, cuda_functions.cu:8-242):

* self mobility of one sphere against the published wall series
  (Swan & Brady 2007; Faxen's parallel and Brenner's normal expansions):
  tt_par = [1 - 9/16 x + 1/8 x^3 - 1/16 x^5]/(6 pi mu a),
  tt_perp = [1 - 9/8 x + 1/2 x^3 - 1/8 x^5]/(6 pi mu a),
  rr_par = [1 - 5/16 x^3]/(8 pi mu a^3), rr_perp = [1 - 1/8 x^3]/(8 pi mu a^3),
  with x = a/h, and a translation-rotation coupling of size
  3/32 x^4/(6 pi mu a^2);
* symmetry of the assembled 6N x 6N matrix (M_ij = M_ji^T) -- which fixes the
  signs of the wall rotation-translation terms relative to each other;
* far from the wall it tends to the unbounded RPY far field;
* positive definiteness for non-overlapping spheres above the wall.
"""
import math

import numpy as np

from oracle import binding as orc


def dense(xs, a, mu):
    n = len(xs)
    M = np.zeros((6 * n, 6 * n))
    for j in range(n):
        for c in range(6):
            F = np.zeros((n, 6))
            F[j, c] = 1.0
            M[:, 6 * j + c] = orc.Mss_wall(xs, F, a, mu).ravel()
    return M


def slab(n, seed, a=1.0):
    rng = np.random.default_rng(seed)
    pts = []
    while len(pts) < n:
        p = rng.uniform([-12, -12, 1.3 * a], [12, 12, 9.0])
        if all(np.linalg.norm(p - q) > 2.2 * a for q in pts):
            pts.append(p)
    return np.array(pts)


def test_self_mobility_matches_wall_series():
    for a, mu, h in ((1.0, 1.0, 1.5), (1.0, 1.0, 4.0), (0.7, 2.0, 3.0)):
        M = dense(np.array([[0.3, -0.2, h]]), a, mu)
        x = a / h
        tt = 1.0 / (6 * math.pi * mu * a)
        rr = 1.0 / (8 * math.pi * mu * a ** 3)
        np.testing.assert_allclose(M[0, 0], tt * (1 - 9 / 16 * x + x ** 3 / 8 - x ** 5 / 16), rtol=1e-13)
        np.testing.assert_allclose(M[1, 1], M[0, 0], rtol=1e-14)
        np.testing.assert_allclose(M[2, 2], tt * (1 - 9 / 8 * x + x ** 3 / 2 - x ** 5 / 8), rtol=1e-13)
        np.testing.assert_allclose(M[3, 3], rr * (1 - 5 / 16 * x ** 3), rtol=1e-13)
        np.testing.assert_allclose(M[5, 5], rr * (1 - x ** 3 / 8), rtol=1e-13)
        c = 3 / 32 * x ** 4 / (6 * math.pi * mu * a ** 2)
        np.testing.assert_allclose(abs(M[0, 4]), c, rtol=1e-13)
        np.testing.assert_allclose(abs(M[1, 3]), c, rtol=1e-13)
        assert M[0, 4] == -M[1, 3]
        assert np.abs(M[2, 3:]).max() == 0 and np.abs(M[0:2, 5]).max() == 0


def test_assembled_matrix_is_symmetric_and_positive_definite():
    xs = slab(24, 1)
    M = dense(xs, 1.0, 1.0)
    assert np.abs(M - M.T).max() <= 1e-14 * np.abs(M).max()
    ev = np.linalg.eigvalsh(0.5 * (M + M.T))
    assert ev.min() > 0


def test_far_from_the_wall_is_unbounded_rpy():
    xs = slab(16, 2)
    xs[:, 2] += 1e7
    rng = np.random.default_rng(3)
    fs = rng.normal(size=(16, 6))
    W = orc.Mss_wall(xs, fs, 1.0, 1.0)
    U = orc.Mss(xs, fs, 1.0, 1.0)   # all pairs farther than 2a: the far-field branch
    assert np.abs(W - U).max() <= 1e-6 * np.abs(U).max()
