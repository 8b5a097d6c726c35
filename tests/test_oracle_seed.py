"""The seeding restatement (oracle/seed.cpp) against known answers.

For N = 4, 6 and 12 the relaxed repulsive seeding must reach the regular
tetrahedron, octahedron and icosahedron (pairwise distances fixed by
geometry), to the relaxation's 1e-4 stopping tolerance.  The two summation
orders the reference uses (its CPU pair loop for N < 1000, its GPU kernel
otherwise) agree to rounding.  Band seedings stay on the unit sphere inside
the band the reference lays out.
"""
import math

import numpy as np

from generic_mobility_interface_class_amd import geometry
from oracle import binding as orc


def pdist(X):
    d = np.linalg.norm(X[:, None] - X[None], axis=-1)
    return np.sort(d[np.triu_indices(len(X), 1)])


def test_platonic_equilibria():
    X, it = orc.seed_uniform(4)
    assert it > 0
    np.testing.assert_allclose(pdist(X), math.sqrt(8.0 / 3.0), atol=2e-3)
    X, _ = orc.seed_uniform(6)
    d = pdist(X)
    np.testing.assert_allclose(d[:12], math.sqrt(2.0), atol=2e-3)
    np.testing.assert_allclose(d[12:], 2.0, atol=2e-3)
    X, _ = orc.seed_uniform(12)
    d = pdist(X)
    edge = 4.0 / math.sqrt(10.0 + 2.0 * math.sqrt(5.0))   # icosahedron on the unit sphere
    np.testing.assert_allclose(d[:30], edge, atol=3e-3)
    np.testing.assert_allclose(d[30:60], edge * (1 + math.sqrt(5.0)) / 2, atol=3e-3)
    np.testing.assert_allclose(d[60:], 2.0, atol=3e-3)


def test_summation_orders_agree_and_points_are_unit():
    a, ia = orc.seed_uniform(300, order=0)
    b, ib = orc.seed_uniform(300, order=1)
    assert ia == ib
    assert np.abs(a - b).max() < 1e-10
    np.testing.assert_allclose(np.linalg.norm(b, axis=1), 1.0, atol=1e-14)
    d = pdist(b)
    assert d.min() > 0.1   # no two points collapse


def test_band_seedings():
    for platy in (False, True):
        X = orc.seed_band(40, platy, 1.0)
        np.testing.assert_allclose(np.linalg.norm(X, axis=1), 1.0, atol=1e-14)
        h = 2.0 * math.sin(0.5)
        nb = round(0.9 * 40) if platy else 40
        assert np.abs(X[:nb, 2]).max() <= 0.5 * h + 2.0 / (math.ceil(nb * 2 / h) - 1)
        if platy:
            np.testing.assert_allclose(X[nb:, 2], -0.9)


def test_seed_files_round_trip(tmp_path):
    X, _ = orc.seed_uniform(50)
    for per_line in (False, True):
        p = tmp_path / f"s{int(per_line)}.seed"
        geometry.write_seed_file(str(p), X, per_line)
        Y = geometry.read_seed_file(str(p), 50)
        np.testing.assert_allclose(Y, X, rtol=1e-5, atol=1e-6)   # "%g": 6 digits
        assert Y.shape == (50, 3)
