"""GPU parity of the device-resident implicit step against the CPU oracle.

Stage by stage first (forces + rigid-body solve, residual, Jacobian inverse
product), then whole steps.  Bars:
  * body mobility reference and rigid solve: the GMRES tolerance (the device
    uses CGS2 Arnoldi, the reference MGS; both stop at 1e-8 relative);
  * residual, J^-1 e: 1e-10 relative;
  * trajectories: equal Broyden iteration counts and states within DL*TOL
    (positions) / TOL (Lie elements) -- the Broyden tolerance.
"""
import numpy as np
import pytest

from generic_mobility_interface_class_amd import capi, geometry
from oracle import binding as orc

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    d = np.abs(b).max()
    return float(np.abs(a - b).max() / d) if d > 0 else float(np.abs(a).max())


def pair(name="step_962", seed=1, **kw):
    cfg = geometry.sim_workload(name, seed=seed, **kw)
    return cfg, orc.Sim(cfg), capi.Sim(cfg)


def test_body_reference_matrices(gpu):
    cfg, o, g = pair()
    print("refmats step_962", flush=True)
    o.make_body_reference_matrices()
    g.make_body_reference_matrices()
    mo, mg = o.mobility(), g.mobility()
    assert rel(mg["KTKinv"], mo["KTKinv"]) < 1e-12
    assert rel(mg["Nref"], mo["Nref"]) < 1e-7


def test_stages_at_first_step(gpu):
    cfg, o, g = pair()
    print("stages step_962", flush=True)
    o.make_body_reference_matrices()
    g.make_body_reference_matrices()
    o.initial_guess(0); o.forces(); o.compute_velocities()
    g.call("gmi_initial_guess", 0)
    g.call("gmi_forces_and_torques", 0)
    its = capi.C.c_int(0)
    g.call("gmi_compute_velocities", 0, capi.C.byref(its))
    mo, mg = o.mobility(), g.mobility()
    print("rigid solve rel err: v_segs %.2e U %.2e lambda %.2e" % (
        rel(mg["v_segs"], mo["v_segs"]), rel(mg["v_bodies"], mo["v_bodies"]),
        rel(mg["f_blobs"], mo["f_blobs"])), flush=True)
    assert rel(mg["v_segs"], mo["v_segs"]) < 1e-7
    assert rel(mg["f_blobs"], mo["f_blobs"]) < 1e-6
    # residual
    eo, big_o = o.errors(0)
    big, bad = capi.C.c_int(0), capi.C.c_int(0)
    g.call("gmi_compute_errors", 0, 0, capi.C.byref(big), capi.C.byref(bad))
    eg = g.broyden(0)
    assert bool(big.value) == big_o and not bad.value
    assert rel(eg, eo) < 1e-6
    # approximate Jacobian inverse (filaments + Schur complement) applied to e
    o.prepare_jacobian(0)
    g.call("gmi_prepare_jacobian_inv", 0)
    g.call("gmi_broyden_find_update")           # iter 0: update = -gamma J^-1 e
    ug = g.broyden(2)
    uo = -cfg.jacobian_confidence * o.jacobian_inv_mult(eg)
    assert rel(ug, uo) < 1e-10


def test_trajectory_ten_steps(gpu):
    cfg, o, g = pair()
    o.make_body_reference_matrices()
    g.make_body_reference_matrices()
    DL = cfg.swimmers[0].DL
    for nt in range(10):
        ro = o.step(nt)
        rg = g.step(nt)
        so, sg = o.state(), g.state()
        dx = np.abs(sg["x"] - so["x"]).max()
        du = np.abs(sg["u"] - so["u"]).max()
        print(f"step {nt}: oracle {ro} gpu {rg} |dx| {dx:.2e} |du| {du:.2e}", flush=True)
        assert ro[0] == 0 and rg[0] == 0
        assert abs(ro[1] - rg[1]) <= 1
        assert dx <= DL * cfg.tol and du <= cfg.tol


def test_fitted_gmres_guess_keeps_the_trajectory(gpu):
    """The history-fitted GMRES initial guess changes where GMRES starts,
    not what it converges to: the trajectory follows the oracle within the
    Broyden tolerance, with far fewer GMRES iterations than the
    previous-solution warm start."""
    cfg, o, g = pair()
    plain = capi.Sim(cfg)
    for s in (o, g, plain):
        s.make_body_reference_matrices()
    g.set_gmres_history(16)
    DL = cfg.swimmers[0].DL
    its_fit = its_plain = 0
    for nt in range(12):
        ro, rg, rp = o.step(nt), g.step(nt), plain.step(nt)
        so, sg = o.state(), g.state()
        dx = np.abs(sg["x"] - so["x"]).max()
        du = np.abs(sg["u"] - so["u"]).max()
        print(f"step {nt}: oracle {ro} fitted {rg} plain {rp} |dx| {dx:.2e} |du| {du:.2e}",
              flush=True)
        assert ro[0] == 0 and rg[0] == 0 and rp[0] == 0
        assert abs(ro[1] - rg[1]) <= 1
        assert dx <= DL * cfg.tol and du <= cfg.tol
        if nt >= 4:
            its_fit += rg[2]
            its_plain += rp[2]
    print(f"GMRES iterations over steps 4-11: fitted {its_fit}, plain {its_plain}", flush=True)
    assert its_fit < 0.7 * its_plain


@pytest.mark.parametrize("name,precision", [("step_962", capi.GMI_FP64),
                                            ("step_9122", capi.GMI_FP64),
                                            ("step_962", capi.GMI_FP32_KAHAN)])
def test_body_inverse_preconditioner_keeps_the_trajectory(gpu, name, precision):
    """The exact body-frame preconditioner (with the preconditioned
    directions kept for the update) changes how fast GMRES converges, not
    what it converges to: trajectories follow the plain FP64 run within the
    Broyden tolerance (the oracle on the small swimmer), also with mixed
    precision, and GMRES takes a few iterations per solve."""
    cfg = geometry.sim_workload(name, seed=1)
    plain, inv = capi.Sim(cfg), capi.Sim(cfg, precision=precision)
    for s in (plain, inv):
        s.make_body_reference_matrices()
    inv.set_body_preconditioner(capi.GMI_PRECOND_BODY_INVERSE)
    o = None
    if name == "step_962":
        o = orc.Sim(cfg)
        o.make_body_reference_matrices()
    DL = cfg.swimmers[0].DL
    for nt in range(6):
        rp, ri = plain.step(nt), inv.step(nt)
        dx = np.abs(inv.state()["x"] - plain.state()["x"]).max()
        print(f"{name} precision {precision} step {nt}: plain {rp} body-inverse {ri} "
              f"|dx| {dx:.2e}", flush=True)
        assert rp[0] == 0 and ri[0] == 0
        assert abs(rp[1] - ri[1]) <= 1
        assert dx <= DL * cfg.tol
        assert ri[2] <= 4 * (ri[1] + 1)  # a few GMRES iterations per solve
        if o is not None:
            ro = o.step(nt)
            assert ro[0] == 0
            assert np.abs(inv.state()["x"] - o.state()["x"]).max() <= DL * cfg.tol


def test_mixed_precision_trajectory(gpu):
    cfg = geometry.sim_workload("step_962", seed=1)
    g64 = capi.Sim(cfg)
    g32 = capi.Sim(cfg, precision=capi.GMI_FP32_KAHAN)
    g32.set_sim_precision(capi.GMI_FP32_KAHAN)
    g64.make_body_reference_matrices()
    g32.make_body_reference_matrices()
    DL = cfg.swimmers[0].DL
    for nt in range(10):
        r64 = g64.step(nt)
        r32 = g32.step(nt)
        s64, s32 = g64.state(), g32.state()
        dx = np.abs(s32["x"] - s64["x"]).max()
        du = np.abs(s32["u"] - s64["u"]).max()
        print(f"mixed step {nt}: fp64 {r64} mixed {r32} |dx| {dx:.2e} |du| {du:.2e}", flush=True)
        assert r64[0] == 0 and r32[0] == 0
        assert dx <= DL * cfg.tol and du <= cfg.tol


def test_uneven_independent_swimmers(gpu):
    sw = [geometry.SwimmerGeometry(nfil=4, nseg=10, nblob=92, R=4.0,
                                   centre=np.array([0.0, 0.0, 0.0])),
          geometry.SwimmerGeometry(nfil=9, nseg=10, nblob=162, R=5.0,
                                   centre=np.array([5.0, 0.0, 0.0])),
          geometry.SwimmerGeometry(nfil=6, nseg=10, nblob=92, R=4.0,
                                   centre=np.array([0.0, 60.0, 0.0]))]
    rng = geometry.Xoshiro256ss(5)
    cfg = geometry.SimConfig(swimmers=sw, phases=2 * np.pi * rng.uniform(19))
    o, g = orc.Sim(cfg), capi.Sim(cfg)
    o.make_body_reference_matrices()
    g.make_body_reference_matrices()
    assert rel(g.mobility()["Nref"], o.mobility()["Nref"]) < 1e-7
    DL = sw[0].DL
    for nt in range(4):
        ro, rg = o.step(nt), g.step(nt)
        so, sg = o.state(), g.state()
        dx = np.abs(sg["x"] - so["x"]).max()
        print(f"uneven step {nt}: oracle {ro} gpu {rg} |dx| {dx:.2e}", flush=True)
        assert ro[0] == 0 and rg[0] == 0
        assert dx <= DL * cfg.tol
        np.testing.assert_allclose(sg["body"], so["body"], rtol=0, atol=DL * cfg.tol)


def test_independent_contexts_step_identically(gpu):
    """With the same inputs and launch shape, two contexts step to the same
    bits (no floating-point atomics; fixed reduction orders), here through
    the mixed-precision path with the body-frame preconditioner."""
    cfg = geometry.sim_workload("step_962", seed=1)
    sims = [capi.Sim(cfg, precision=capi.GMI_FP32_KAHAN) for _ in range(2)]
    for s in sims:
        s.call("gmi_set_launch_shape", 4096, 128)
        s.set_body_preconditioner(capi.GMI_PRECOND_BODY_INVERSE)
        s.make_body_reference_matrices()
    for nt in range(4):
        r = [s.step(nt) for s in sims]
        assert r[0] == r[1] and r[0][0] == 0
        st = [s.state() for s in sims]
        assert np.array_equal(st[0]["x"], st[1]["x"]) and np.array_equal(st[0]["u"], st[1]["u"])


def test_graph_replay_matches_eager(gpu):
    """The captured Arnoldi chunks are replayed across many solves with host
    work in between; every replay must equal the eager launches bit for bit."""
    cfg = geometry.sim_workload("step_962", seed=1)
    ge = capi.Sim(cfg)
    ge.set_graphs(False)
    gg = capi.Sim(cfg)
    gg.set_graphs(True, chunk=4)
    ge.make_body_reference_matrices()
    gg.make_body_reference_matrices()
    for nt in range(5):
        re_, rg = ge.step(nt), gg.step(nt)
        assert re_ == rg
        se, sg = ge.state(), gg.state()
        assert np.array_equal(se["x"], sg["x"]) and np.array_equal(se["u"], sg["u"])
    me, mg = ge.mobility(), gg.mobility()
    assert np.array_equal(me["f_blobs"], mg["f_blobs"])
    # FP64 GMRES without the fitted guess: every Broyden iteration after the
    # first residual ran from a captured graph (its solve finished eagerly
    # past the graph's first chunk)
    assert ge.graph_stats()[0] == 0 and gg.graph_stats()[0] > 0


@pytest.mark.parametrize("precision,fitted", [(capi.GMI_FP64, False), (capi.GMI_FP32_KAHAN, False),
                                              (capi.GMI_FP64, True)])
def test_captured_broyden_iterations_match_eager(gpu, precision, fitted):
    """Under the exact body-frame preconditioner (FP64 Krylov products, one
    Arnoldi step per check) each Broyden iteration is one captured graph, and
    so it is with the reference preconditioner in FP64 with the history-fitted
    GMRES guess (its fit and push run inside the graph, the history's count
    and head kept on the device).  The kernels are the eager ones in the same
    order, so iteration counts, trajectory and solver state equal the eager
    run's bit for bit, including the solves that need more than the graph's
    first Arnoldi chunk (finished eagerly) and, in mixed precision, the FP32
    pair sums of the velocities."""
    cfg = geometry.sim_workload("step_962", seed=1)
    sims = []
    for graphs in (False, True):
        g = capi.Sim(cfg)
        g.set_sim_precision(precision)
        g.set_graphs(graphs)
        if fitted:
            g.set_gmres_history(16)
        else:
            g.set_body_preconditioner(capi.GMI_PRECOND_BODY_INVERSE)
        g.make_body_reference_matrices()
        sims.append(g)
    ge, gg = sims
    for nt in range(8):
        re_, rg = ge.step(nt), gg.step(nt)
        print(f"step {nt}: eager {re_} captured {rg} graph stats {gg.graph_stats()}", flush=True)
        assert re_ == rg and re_[0] == 0
        se, sg = ge.state(), gg.state()
        for k in ("x", "u", "lam"):
            assert np.array_equal(se[k], sg[k]), k
    assert np.array_equal(ge.mobility()["f_blobs"], gg.mobility()["f_blobs"])
    assert np.array_equal(ge.broyden(2), gg.broyden(2))
    n_it, n_graphs = gg.graph_stats()
    assert ge.graph_stats()[0] == 0 and n_it > 0 and n_graphs > 0


@pytest.mark.parametrize("passes", [1, 2])
def test_gram_schmidt_profile_counts_the_arnoldi_steps(gpu, passes):
    """gmi_profile_krylov: eagerly launched Arnoldi steps are bracketed one
    each; steps launched past convergence inside a chunk (flag-skipped) are
    not counted, so the count is the GMRES iteration total of the step, and
    the bytes are 8 p n (k + 2) summed over steps k = 1..its of each solve
    (p Gram-Schmidt passes, n the padded GMRES length).
    Captured graphs record nothing."""
    cfg = geometry.sim_workload("step_962", seed=1)
    g = capi.Sim(cfg)
    g.set_gram_schmidt(passes)
    g.make_body_reference_matrices()
    n = 3 * cfg.swimmers[0].nblob + 6
    for use_graphs in (False, True):
        g.set_graphs(use_graphs, chunk=6)
        g.call("gmi_profile_begin")
        rc, _, gits = g.step(int(use_graphs))
        assert rc == 0 and gits > 0
        ms, nbytes, steps = capi.C.c_double(0.0), capi.C.c_double(0.0), capi.C.c_int(0)
        g.call("gmi_profile_krylov", capi.C.byref(ms), capi.C.byref(nbytes), capi.C.byref(steps))
        pm, pn = capi.C.c_double(0.0), capi.C.c_int(0)
        g.call("gmi_profile_end", capi.C.byref(pm), capi.C.byref(pn))
        print(f"graphs {use_graphs}: GMRES {gits}, recorded {steps.value} steps, "
              f"{ms.value:.3f} ms, {nbytes.value:.3e} B", flush=True)
        if use_graphs:
            assert steps.value == 0 and nbytes.value == 0.0
            continue
        assert steps.value == gits
        assert ms.value > 0
        # 8 p n' (k + 2) per step with n' >= n (the padded length) and k >= 1
        assert nbytes.value >= 8.0 * passes * n * 3 * gits


@pytest.mark.parametrize("nseg", [23, 30, 48, 64])
def test_long_filaments_match_oracle(gpu, nseg):
    """Filaments longer than 21 segments: NSEG = 23 (6 NSEG = 138 columns,
    the Gauss-Jordan inverse held in LDS), NSEG = 30, 48 and 64 (up to 384
    columns, the matrix in global memory; 64 is the largest NSEG a time step
    takes).  J^-1 e at the first step and four steps against the oracle."""
    cfg = geometry.sim_config(nfil=6, nseg=nseg, nblob=162, R=6.0, seed=2)
    o, g = orc.Sim(cfg), capi.Sim(cfg)
    print(f"long filaments nseg {nseg}", flush=True)
    o.make_body_reference_matrices()
    g.make_body_reference_matrices()
    o.initial_guess(0); o.forces(); o.compute_velocities()
    g.call("gmi_initial_guess", 0)
    g.call("gmi_forces_and_torques", 0)
    its = capi.C.c_int(0)
    g.call("gmi_compute_velocities", 0, capi.C.byref(its))
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
    cfg2 = geometry.sim_config(nfil=6, nseg=nseg, nblob=162, R=6.0, seed=2)
    o, g = orc.Sim(cfg2), capi.Sim(cfg2)
    o.make_body_reference_matrices()
    g.make_body_reference_matrices()
    DL = cfg2.swimmers[0].DL
    for nt in range(4):
        ro, rg = o.step(nt), g.step(nt)
        so, sg = o.state(), g.state()
        dx = np.abs(sg["x"] - so["x"]).max()
        du = np.abs(sg["u"] - so["u"]).max()
        print(f"  step {nt}: oracle {ro} gpu {rg} |dx| {dx:.2e} |du| {du:.2e}", flush=True)
        assert ro[0] == 0 and rg[0] == 0
        assert abs(ro[1] - rg[1]) <= 1
        assert dx <= DL * cfg2.tol and du <= cfg2.tol


_ORACLE_9122 = {}


def _oracle_9122_steps(cfg, nsteps):
    """States of the oracle's first `nsteps` steps of the 9,122-particle
    swimmer (a few seconds per step on 16 host threads), once per session."""
    if not _ORACLE_9122:
        o = orc.Sim(cfg)
        o.make_body_reference_matrices()
        out = []
        for nt in range(nsteps):
            r = o.step(nt)
            print(f"  oracle step_9122 step {nt}: {r}", flush=True)
            out.append((r, o.state()))
        _ORACLE_9122["steps"] = out
    return _ORACLE_9122["steps"]


@pytest.mark.parametrize("precision, passes", [(capi.GMI_FP64, 1), (capi.GMI_FP32_KAHAN, 1),
                                               (capi.GMI_FP64, 2)])
def test_step_9122_matches_oracle(gpu, precision, passes):
    """The 5,120-segment swimmer (256 x 20 segments + 4,002 blobs), three
    steps on the GPU against the CPU oracle, FP64 and mixed precision, with
    one Gram-Schmidt pass per Arnoldi step (as the oracle's modified
    Gram-Schmidt) and with two: Broyden counts within one, positions within
    DL*TOL, Lie elements within TOL."""
    cfg = geometry.sim_workload("step_9122", seed=1)
    print(f"step_9122 precision {precision} passes {passes}", flush=True)
    ref = _oracle_9122_steps(cfg, 3)
    g = capi.Sim(cfg, precision=precision)
    g.set_gram_schmidt(passes)
    if precision == capi.GMI_FP32_KAHAN:
        g.set_sim_precision(capi.GMI_FP32_KAHAN)
    g.make_body_reference_matrices()
    DL = cfg.swimmers[0].DL
    for nt, (ro, so) in enumerate(ref):
        rg = g.step(nt)
        sg = g.state()
        dx = np.abs(sg["x"] - so["x"]).max()
        du = np.abs(sg["u"] - so["u"]).max()
        print(f"  step {nt}: oracle {ro} gpu {rg} |dx| {dx:.2e} |du| {du:.2e}", flush=True)
        assert ro[0] == 0 and rg[0] == 0
        assert abs(ro[1] - rg[1]) <= 1
        assert dx <= DL * cfg.tol and du <= cfg.tol


@pytest.mark.parametrize("precond", [capi.GMI_PRECOND_SELF_MOBILITY, capi.GMI_PRECOND_BODY_INVERSE])
def test_coupled_swimmers_match_oracle(gpu, precond):
    """Two coupled swimmers (every particle couples to every other, the
    reference's NSWIM > 1 run) stepped on the GPU against the oracle: the
    work lists keep target tiles inside one swimmer and give them every
    source; the per-swimmer preconditioners stay preconditioners."""
    sw = [geometry.SwimmerGeometry(nfil=4, nseg=10, nblob=92, R=4.0,
                                   centre=np.array([0.0, 0.0, 0.0])),
          geometry.SwimmerGeometry(nfil=6, nseg=10, nblob=92, R=4.0,
                                   centre=np.array([60.0, 0.0, 0.0]))]
    rng = geometry.Xoshiro256ss(5)
    cfg = geometry.SimConfig(swimmers=sw, phases=2 * np.pi * rng.uniform(10), interact=True)
    o, g = orc.Sim(cfg), capi.Sim(cfg)
    print(f"coupled swimmers preconditioner {precond}", flush=True)
    o.make_body_reference_matrices()
    g.make_body_reference_matrices()
    g.set_body_preconditioner(precond)
    DL = sw[0].DL
    for nt in range(5):
        ro, rg = o.step(nt), g.step(nt)
        so, sg = o.state(), g.state()
        dx = np.abs(sg["x"] - so["x"]).max()
        print(f"  step {nt}: oracle {ro} gpu {rg} |dx| {dx:.2e}", flush=True)
        assert ro[0] == 0 and rg[0] == 0
        assert abs(ro[1] - rg[1]) <= 1
        assert dx <= DL * cfg.tol
        np.testing.assert_allclose(sg["body"], so["body"], rtol=0, atol=DL * cfg.tol)


@pytest.mark.parametrize("always_pivot", [False, True])
def test_jacobian_inverse_modes_match_oracle(gpu, always_pivot):
    """The filament inverses of the 962-particle swimmer (6 NSEG = 120), by
    the checked pivot-free Gauss-Jordan (default) and by partial pivoting
    for every filament: J^-1 e at the first step against the oracle's LU
    inverse, then five steps with the oracle's Broyden counts."""
    cfg = geometry.sim_workload("step_962", seed=1)
    o, g = orc.Sim(cfg), capi.Sim(cfg)
    g.set_jacobian_pivoting(always_pivot)
    print(f"jacobian inverse always_pivot {always_pivot}", flush=True)
    o.make_body_reference_matrices()
    g.make_body_reference_matrices()
    o.initial_guess(0); o.forces(); o.compute_velocities()
    g.call("gmi_initial_guess", 0)
    g.call("gmi_forces_and_torques", 0)
    its = capi.C.c_int(0)
    g.call("gmi_compute_velocities", 0, capi.C.byref(its))
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
    o, g = orc.Sim(cfg), capi.Sim(cfg)
    g.set_jacobian_pivoting(always_pivot)
    o.make_body_reference_matrices()
    g.make_body_reference_matrices()
    DL = cfg.swimmers[0].DL
    for nt in range(5):
        ro, rg = o.step(nt), g.step(nt)
        dx = np.abs(g.state()["x"] - o.state()["x"]).max()
        print(f"  step {nt}: oracle {ro} gpu {rg} |dx| {dx:.2e}", flush=True)
        assert ro[0] == 0 and rg[0] == 0
        assert abs(ro[1] - rg[1]) <= 1
        assert dx <= DL * cfg.tol
