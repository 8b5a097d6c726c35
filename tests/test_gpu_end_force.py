"""The reference's default physics on the device step: instability-driven
cilia (CILIA_TYPE 0) with a follower force END_FORCE_MAGNITUDE at every
filament tip, the nt = 0 perturbations on segments 4 and 12, and the end
force's approximate-Jacobian and Schur-complement terms (filament.cpp:
1177-1202,1466-1470; swimmer.cpp:497-503,557-571,630-633), against the
oracle.  Bars as for the other step tests: J^-1 e to 1e-10, Broyden counts
within one, positions within DL*TOL and Lie elements within TOL.
"""
import numpy as np
import pytest

from generic_mobility_interface_class_amd import capi, geometry
from oracle import binding as orc

pytestmark = pytest.mark.gpu


def rel(a, b):
    d = np.abs(b).max()
    return float(np.abs(a - b).max() / d)


def _steps(cfg, nsteps, precision=capi.GMI_FP64, label=""):
    o, g = orc.Sim(cfg), capi.Sim(cfg, precision=precision)
    if precision == capi.GMI_FP32_KAHAN:
        g.set_sim_precision(capi.GMI_FP32_KAHAN)
    o.make_body_reference_matrices()
    g.make_body_reference_matrices()
    DL = cfg.swimmers[0].DL
    for nt in range(nsteps):
        ro, rg = o.step(nt), g.step(nt)
        so, sg = o.state(), g.state()
        dx = np.abs(sg["x"] - so["x"]).max()
        du = np.abs(sg["u"] - so["u"]).max()
        print(f"{label} step {nt}: oracle {ro} gpu {rg} |dx| {dx:.2e} |du| {du:.2e}", flush=True)
        assert ro[0] == 0 and rg[0] == 0
        assert abs(ro[1] - rg[1]) <= 1
        assert dx <= DL * cfg.tol and du <= cfg.tol


def test_end_force_stages_match_oracle(gpu):
    g0 = geometry.SwimmerGeometry(nfil=6, nseg=20, nblob=162, R=6.0)
    cfg = geometry.instability_config([g0], seed=7)
    print("end force stages", flush=True)
    o, g = orc.Sim(cfg), capi.Sim(cfg)
    o.make_body_reference_matrices()
    g.make_body_reference_matrices()
    o.initial_guess(0); o.forces(); o.compute_velocities()
    g.call("gmi_initial_guess", 0)
    g.call("gmi_forces_and_torques", 0)
    its = capi.C.c_int(0)
    g.call("gmi_compute_velocities", 0, capi.C.byref(its))
    mo, mg = o.mobility(), g.mobility()
    # velocities of the forces with the end force and the perturbations
    assert rel(mg["v_segs"], mo["v_segs"]) < 1e-7
    big, bad = capi.C.c_int(0), capi.C.c_int(0)
    g.call("gmi_compute_errors", 0, 0, capi.C.byref(big), capi.C.byref(bad))
    eg = g.broyden(0)
    o.prepare_jacobian(0)
    g.call("gmi_prepare_jacobian_inv", 0)
    g.call("gmi_broyden_find_update")
    ug = g.broyden(2)
    uo = -cfg.jacobian_confidence * o.jacobian_inv_mult(eg)
    print(f"  J^-1 e rel err {rel(ug, uo):.2e}", flush=True)
    assert rel(ug, uo) < 1e-10


@pytest.mark.parametrize("precision", [capi.GMI_FP64, capi.GMI_FP32_KAHAN])
def test_end_force_swimmer_matches_oracle(gpu, precision):
    g0 = geometry.SwimmerGeometry(nfil=6, nseg=20, nblob=162, R=6.0)
    cfg = geometry.instability_config([g0], seed=7)
    _steps(cfg, 8, precision, f"end force precision {precision}")


@pytest.mark.parametrize("nfil", [1, 16])
def test_reference_default_swimmer_matches_oracle(gpu, tmp_path, nfil):
    """The reference's default run (config.hpp): NSEG 20, NBLOB 790 blobs of
    radius RBLOB 1.4 seeded as the reference seeds them on a sphere of
    radius L, NFIL 1 (and 16), END_FORCE_MAGNITUDE 220 KB/(DL NSEG)^2,
    DT 36.3833/300, GMRES tolerance 1e-10: ten steps."""
    ctx = capi.Context([1], [1], 1)
    g0 = geometry.reference_swimmer(ctx, nfil=nfil, nseg=20, nblob=790,
                                    cache_dir=str(tmp_path))
    cfg = geometry.instability_config([g0], seed=11)
    _steps(cfg, 10, capi.GMI_FP64, f"reference default nfil {nfil}")
