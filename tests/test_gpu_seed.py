"""Reference seeding on the GPU (gmi_seed_sphere) against the oracle, and a
swimmer laid out as the reference lays one out (uniform-seeded blobs of
radius RBLOB = 1.4 on a sphere of radius L, config.hpp:102-103; 790 blobs,
the reference's NBLOB) stepped on the GPU against the oracle."""
import numpy as np
import pytest

from generic_mobility_interface_class_amd import capi, geometry
from oracle import binding as orc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [12, 790, 1500])
def test_uniform_seeding_matches_oracle(gpu, n):
    ctx = capi.Context([1], [1], 1)
    X, it = ctx.seed_sphere(n)
    Y, ito = orc.seed_uniform(n, order=1)
    print(f"seed N={n}: gpu {it} oracle {ito} iterations, max diff {np.abs(X - Y).max():.2e}",
          flush=True)
    assert it == ito
    assert np.abs(X - Y).max() < 1e-9
    np.testing.assert_allclose(np.linalg.norm(X, axis=1), 1.0, atol=1e-14)


def test_band_seedings_match_oracle(gpu):
    ctx = capi.Context([1], [1], 1)
    for kind, platy in ((capi.GMI_SEED_EQUATORIAL, False), (capi.GMI_SEED_PLATY, True)):
        X, _ = ctx.seed_sphere(64, kind, 1.0)
        assert np.abs(X - orc.seed_band(64, platy, 1.0)).max() < 1e-15


def test_reference_layout_swimmer_steps_match_oracle(gpu, tmp_path):
    ctx = capi.Context([1], [1], 1)
    g = geometry.reference_swimmer(ctx, nfil=16, nseg=20, nblob=790, cache_dir=str(tmp_path))
    assert (tmp_path / "sphere790.seed").exists() and (tmp_path / "sphere16.seed").exists()
    # the cached files are read back (6 digits), as the reference does on its next run
    g2 = geometry.reference_swimmer(None, nfil=16, nseg=20, nblob=790, cache_dir=str(tmp_path))
    assert np.abs(g2.blob_references() - g.blob_references()).max() < 1e-4 * g.R
    rng = geometry.Xoshiro256ss(3)
    cfg = geometry.SimConfig(swimmers=[g2], phases=2 * np.pi * rng.uniform(16))
    o, s = orc.Sim(cfg), capi.Sim(cfg)
    o.make_body_reference_matrices()
    s.make_body_reference_matrices()
    assert np.abs(s.mobility()["Nref"] - o.mobility()["Nref"]).max() <= \
        1e-7 * np.abs(o.mobility()["Nref"]).max()
    DL = g2.DL
    for nt in range(3):
        ro, rg = o.step(nt), s.step(nt)
        dx = np.abs(s.state()["x"] - o.state()["x"]).max()
        print(f"reference layout step {nt}: oracle {ro} gpu {rg} |dx| {dx:.2e}", flush=True)
        assert ro[0] == 0 and rg[0] == 0 and abs(ro[1] - rg[1]) <= 1
        assert dx <= DL * cfg.tol
