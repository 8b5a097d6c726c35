"""Wall-bounded segment mobility on the GPU (gmi_set_wall, k_pair_wall)
against the oracle (orc_Mss_wall, pinned in test_oracle_wall.py)."""
import numpy as np
import pytest

from generic_mobility_interface_class_amd import capi
from oracle import binding as orc

pytestmark = pytest.mark.gpu


def carpet(nfil, nseg, seed):
    """Straight filaments standing on the wall plane, tilted at random."""
    rng = np.random.default_rng(seed)
    side = int(np.ceil(np.sqrt(nfil)))
    xs = []
    for f in range(nfil):
        base = np.array([4.0 * (f % side), 4.0 * (f // side), 1.2])
        tilt = rng.normal(scale=0.3, size=2)
        d = np.array([tilt[0], tilt[1], 1.0])
        d /= np.linalg.norm(d)
        for k in range(nseg):
            xs.append(base + 2.2 * k * d)
    return np.array(xs)


@pytest.mark.parametrize("shape", [None, (4096, 32)])
def test_wall_apply_matches_oracle(gpu, shape):
    nfil, nseg = 36, 20
    xs = carpet(nfil, nseg, 1)
    fs = np.random.default_rng(2).normal(size=(len(xs), 6))
    ctx = capi.Context([len(xs)], [0], nseg)
    ctx.set_wall(True)
    if shape:
        ctx.call("gmi_set_launch_shape", *shape)
    ctx.upload(xs, fs)
    ctx.call("gmi_evaluate_segment_segment_mobility")
    vs, _ = ctx.velocities()
    ref = orc.Mss_wall(xs, fs, 1.0, 1.0)
    err = np.abs(vs - ref).max() / np.abs(ref).max()
    print(f"wall apply {len(xs)} segments, shape {shape}: rel err {err:.2e}", flush=True)
    assert err <= 1e-12


def test_wall_refusals(gpu):
    ctx = capi.Context([20], [12], 20)
    with pytest.raises(capi.GmiError):
        ctx.set_wall(True)
